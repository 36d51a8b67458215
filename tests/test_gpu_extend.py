"""The interface Krylov extension loop run as one launch (every task runs its
own begin / transfer sweep / end steps until it converges) against the
host-driven loop of separate launches: the same extension counts and the same
values bit for bit, for 2D triangles and 3D tetrahedra, windowed and all-HBM
sweeps, several launch shapes, and a cap on the extensions that leaves tasks
unconverged (both forms warn and stop at the cap)."""
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case2d(n, seed, M=12, nphi=24):
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_cell_2d(10.0, 0.25, n, seed=seed)
    mat = synthetic_material_2d(M, nphi, 1.0, 1e4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    return geo, mat, bcs, eff


def _case3d():
    from openbte_amd.geometry import pore_slab_3d
    from openbte_amd.material import synthetic_material_3d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_slab_3d(10.0, 4.0, 0.2, 2000, seed=4, nz=4)
    mat = synthetic_material_3d(8, 4, 6, 1.0, 1e4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    return geo, mat, bcs, eff


def _run(op, sched, cfg, t0, n_iter, max_extend=None):
    import torch
    from openbte_amd.bte import BTEDevice, InterfaceNotConverged
    dev = BTEDevice(op, sched=sched, cfg=cfg)
    dev.keep_solutions(True)
    dev.reset(t0)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter('always', InterfaceNotConverged)
        exts = [dev.outer_iteration(max_extend=max_extend) for _ in range(n_iter)]
    torch.cuda.synchronize()
    n_warn = sum(issubclass(x.category, InterfaceNotConverged) for x in w)
    g = dev.gbuf
    out = {'t': dev.buf['t'].cpu().numpy().copy(), 'kappa': dev.buf['kappa'].cpu().numpy().copy(),
           'j': dev.buf['j'].cpu().numpy().copy(), 'dt': dev.buf['dt'][dev.cur].cpu().numpy().copy(),
           'tb': dev.buf['tb'][dev.cur].cpu().numpy().copy()}
    for k in ('y', 'r', 'k', 'active', 'lane_active', 'rnorm'):
        out['gcr_' + k] = g[k].cpu().numpy().copy()
    del dev
    return out, exts, n_warn


def _compare(op, sc, kw, t0, n_iter=4, max_extend=None):
    from openbte_amd.bte import SweepConfig
    a, ea, wa = _run(op, sc, SweepConfig(fused_extend=True, **kw), t0, n_iter, max_extend)
    b, eb, wb = _run(op, sc, SweepConfig(fused_extend=False, **kw), t0, n_iter, max_extend)
    assert ea == eb, (ea, eb)
    assert wa == wb
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    return ea, wa


@pytest.mark.parametrize('mc,block,window', [(8, 256, True), (4, 128, True), (8, 256, False), (16, 256, True)])
def test_fused_extension_loop_2d(mc, block, window):
    from openbte_amd.operator import build_host_operator, build_schedule
    op = build_host_operator(*_case2d(3000, seed=5))
    sc = build_schedule(op)
    print(f'# fused extend 2d mc={mc} block={block} window={window}', flush=True)
    exts, _ = _compare(op, sc, dict(mc=mc, block=block, window=window), np.linspace(-0.5, 0.5, op.n))
    assert exts[0] > 1


def test_fused_extension_loop_3d():
    from openbte_amd.operator import build_host_operator, build_schedule
    op = build_host_operator(*_case3d())
    sc = build_schedule(op)
    print('# fused extend 3d', op.n, flush=True)
    exts, _ = _compare(op, sc, dict(mc=8, block=256), np.linspace(-0.5, 0.5, op.n))
    assert exts[0] > 1


def test_fused_extension_loop_stops_at_the_cap():
    from openbte_amd.operator import build_host_operator, build_schedule
    op = build_host_operator(*_case2d(2000, seed=6))
    sc = build_schedule(op)
    print('# fused extend cap', flush=True)
    exts, warned = _compare(op, sc, dict(mc=8, block=256), np.linspace(-0.5, 0.5, op.n), n_iter=2, max_extend=2)
    assert exts[0] == 2 and warned >= 1
