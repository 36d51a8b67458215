"""The sweep with an LDS window of recent levels against the sweep that reads
every upwind value from HBM: the same values bit for bit, for the plain and
batched launches, 2D triangles and 3D tetrahedra, transfer sweeps that skip
converged bins, and levels wider than one pass of the workgroup.  The windowed
final sweep forms the mode sums from its window (per-chunk sums); the
all-HBM path forms them from t: the lattice temperatures, fluxes, wall
temperatures and kappa agree bit for bit too, with and without the windowed
sweep storing every value."""
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


def _iterate(op, sched, cfg, t0, n_iter=3, keep=True):
    import torch
    from openbte_amd.bte import BTEDevice
    dev = BTEDevice(op, sched=sched, cfg=cfg)
    dev.keep_solutions(keep)
    dev.reset(t0)
    exts = [dev.outer_iteration() for _ in range(n_iter)]
    torch.cuda.synchronize()
    out = {'kappa': dev.buf['kappa'].cpu().numpy().copy(), 'j': dev.buf['j'].cpu().numpy().copy()}
    if keep:
        out['t'] = _rows(dev)
    for k in ('dt', 'tb'):
        out[k] = dev.buf[k][dev.cur].cpu().numpy().copy()
    cap = dev.window_positions()
    del dev
    return out, exts, cap


def _rows(dev):
    """The solution workspace."""
    return dev.buf['t'].cpu().numpy().copy()


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def _sums_only(a):
    return {k: v for k, v in a.items() if k != 't'}


@pytest.mark.parametrize('mc,block,max_width', [(8, 256, 192), (4, 256, 192), (8, 64, 192), (4, 128, 0)])
def test_window_equals_global_2d(mc, block, max_width):
    from openbte_amd.bte import SweepConfig
    from openbte_amd.operator import build_host_operator, build_schedule
    geo, mat, bcs, eff = _case2d(3000, seed=5)
    op = build_host_operator(geo, mat, bcs, eff)
    sc = build_schedule(op, max_width=max_width)
    t0 = np.linspace(-0.5, 0.5, op.n)
    print(f'# window 2d mc={mc} block={block} max_width={max_width}', flush=True)
    a, ea, cap_a = _iterate(op, sc, SweepConfig(mc=mc, block=block, window=True), t0)
    b, eb, cap_b = _iterate(op, sc, SweepConfig(mc=mc, block=block, window=False), t0)
    c, ec, _ = _iterate(op, sc, SweepConfig(mc=mc, block=block, window=True), t0, keep=False)
    assert cap_a > 0 and cap_b == 0
    assert ea == eb == ec
    _same(a, b)
    _same(c, _sums_only(b))


@pytest.mark.parametrize('max_width', [192, 0])
def test_window_equals_global_3d(max_width):
    from openbte_amd.bte import SweepConfig
    from openbte_amd.geometry import pore_slab_3d
    from openbte_amd.material import synthetic_material_3d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    from openbte_amd.operator import build_host_operator, build_schedule
    geo = pore_slab_3d(10.0, 4.0, 0.2, 2000, seed=4, nz=4)
    mat = synthetic_material_3d(8, 4, 6, 1.0, 1e4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    op = build_host_operator(geo, mat, bcs, eff)
    sc = build_schedule(op, max_width=max_width)  # 0: levels wider than one pass
    t0 = np.linspace(-0.5, 0.5, op.n)
    print('# window 3d', op.n, max_width, flush=True)
    a, ea, cap_a = _iterate(op, sc, SweepConfig(mc=8, block=256, window=True), t0)
    b, eb, cap_b = _iterate(op, sc, SweepConfig(mc=8, block=256, window=False), t0)
    c, ec, _ = _iterate(op, sc, SweepConfig(mc=8, block=256, window=True), t0, keep=False)
    assert cap_a > 0 and cap_b == 0
    assert ea == eb == ec
    _same(a, b)


def test_window_equals_global_batch():
    import torch
    from openbte_amd.batch import BTEBatch
    from openbte_amd.bte import SweepConfig
    from openbte_amd.operator import build_host_operator, build_schedule
    cases = [_case2d(n, seed=s) for n, s in ((800, 1), (2500, 2), (1500, 3))]
    ops = [build_host_operator(*c) for c in cases]
    scheds = [build_schedule(o) for o in ops]
    t0s = [np.linspace(-0.5, 0.5, o.n) for o in ops]
    res = []
    for window, keep in ((True, True), (False, True), (True, False)):
        print(f'# window batch {window} keep {keep}', flush=True)
        b = BTEBatch(ops, cfg=SweepConfig(mc=8, block=256, window=window), scheds=scheds)
        b.keep_solutions(keep)
        b.reset(t0s)
        exts = [b.outer_iteration() for _ in range(3)]
        torch.cuda.synchronize()
        sums = [d.buf['kappa'].cpu().numpy().copy() for d in b.devs]
        sums += [d.buf['dt'][0].cpu().numpy().copy() for d in b.devs]
        sums += [d.buf['j'].cpu().numpy().copy() for d in b.devs]
        res.append(([_rows(d) for d in b.devs] if keep else [], sums, exts))
        del b
    (ta, ka, ea), (tb, kb, eb), (_, kc, ec) = res
    assert ea == eb == ec
    for x, y in zip(ta + ka, tb + kb):
        assert np.array_equal(x, y)
    for x, y in zip(kc, kb):
        assert np.array_equal(x, y)
