"""Batched solves: every problem of a batch gets the result of its own
standalone solve (same arithmetic, per-problem stop rule), and matches the
oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed, n, M=6, nphi=8):
    from openbte_amd.geometry import random_pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = random_pore_cell_2d(seed, n)
    mat = synthetic_material_2d(M, nphi, 1.0, 1.0e4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-geo.size[0] / geo.size[1])
    return dict(geo=geo, mat=mat, bcs=bcs, effective_thermal_conductivity=eff)


def test_batch_equals_single_and_oracle():
    from oracle import bte_oracle as bo, fourier_oracle as fo
    from openbte_amd.bte import BTE_RTA
    from openbte_amd.batch import BTE_RTA_batch
    probs = [_problem(11, 300), _problem(12, 900), _problem(13, 500)]
    for pr in probs:
        pr['fourier'] = fo.fourier(pr['geo'], pr['mat'].thermal_conductivity, pr['bcs'],
                                   pr['effective_thermal_conductivity'])
    outs = BTE_RTA_batch(probs, bte_max_iter=8, autotune=False)
    for pr, out in zip(probs, outs):
        single = BTE_RTA(pr['geo'], pr['mat'], pr['bcs'], fourier=pr['fourier'],
                         effective_thermal_conductivity=pr['effective_thermal_conductivity'], bte_max_iter=8)
        assert out.aux['n_iter'] == single.aux['n_iter']
        assert out.kappa_eff == single.kappa_eff
        np.testing.assert_array_equal(out.variables['Temperature_BTE']['data'],
                                      single.variables['Temperature_BTE']['data'])
        ref = bo.bte_rta(pr['geo'], pr['mat'], pr['bcs'], fourier=pr['fourier'],
                         effective_thermal_conductivity=pr['effective_thermal_conductivity'], bte_max_iter=8)
        assert out.aux['n_iter'] == ref.aux['n_iter']
        assert abs(out.kappa_eff / ref.kappa_eff - 1) < 1e-6
        assert np.abs(out.variables['Temperature_BTE']['data']
                      - ref.variables['Temperature_BTE']['data']).max() < 1e-8


def test_batch_mixed_boundary_kinds(isothermal_case, two_wall_case):
    """One batch holding an isothermal-contact problem with a heat source,
    quads with two-wall cells and a periodic pore cell: each result is bitwise
    its standalone solve."""
    from oracle import fourier_oracle as fo
    from openbte_amd.bte import BTE_RTA
    from openbte_amd.batch import BTE_RTA_batch
    geo, mat, bcs, eff = isothermal_case
    x = geo.elem_centroids
    iso = dict(geo=geo, mat=mat, bcs=bcs, effective_thermal_conductivity=eff,
               heat_source=0.1 / mat.heat_source_coeff * (1.2 + np.sin(0.6 * x[:, 0])))
    g2, m2, b2, e2, _ = two_wall_case
    probs = [iso, dict(geo=g2, mat=m2, bcs=b2, effective_thermal_conductivity=e2), _problem(14, 400)]
    for pr in probs:
        pr['fourier'] = fo.fourier(pr['geo'], pr['mat'].thermal_conductivity, pr['bcs'],
                                   pr['effective_thermal_conductivity'])
    outs = BTE_RTA_batch(probs, bte_max_iter=6, autotune=False)
    for pr, out in zip(probs, outs):
        single = BTE_RTA(pr['geo'], pr['mat'], pr['bcs'], fourier=pr['fourier'],
                         effective_thermal_conductivity=pr['effective_thermal_conductivity'],
                         heat_source=pr.get('heat_source'), bte_max_iter=6)
        assert out.aux['n_iter'] == single.aux['n_iter']
        assert out.kappa_eff == single.kappa_eff
        np.testing.assert_array_equal(out.variables['Temperature_BTE']['data'],
                                      single.variables['Temperature_BTE']['data'])
        np.testing.assert_array_equal(out.variables['Flux_BTE']['data'], single.variables['Flux_BTE']['data'])


@pytest.mark.parametrize('mc', [8, 4])
def test_batch_wave_tasks_beside_workgroup_tasks(mc):
    """A problem whose adjacent levels exceed the wave window (8000 cells, 152
    cells at mc 8) runs one task per workgroup; the narrow ones run one task
    per wave in the same launches (at mc 4 all three do).  Each result is
    bitwise its standalone solve, and the same with wave tasks switched off."""
    from openbte_amd.bte import BTE_RTA, SweepConfig
    from openbte_amd.batch import BTEBatch, BTE_RTA_batch
    from openbte_amd.operator import build_host_operator
    probs = [_problem(41, 8000), _problem(42, 300), _problem(43, 900)]
    ops = [build_host_operator(p['geo'], p['mat'], p['bcs'], p['effective_thermal_conductivity']) for p in probs]
    b = BTEBatch(ops, cfg=SweepConfig(mc=mc, block=256))
    expect = [mc == 4, True, True]
    assert b.wave_problem.tolist() == expect, (b.wave_pair_limit(), [d.lvl_pair_max for d in b.devs])
    assert b.n_wave_tasks > 0 and (b.n_block_tasks > 0) == (not all(expect))
    del b
    outs = BTE_RTA_batch(probs, bte_max_iter=4, config=SweepConfig(mc=mc, block=256))
    blocks = BTE_RTA_batch(probs, bte_max_iter=4, config=SweepConfig(mc=mc, block=256, wave_tasks=False))
    for pr, out, blk in zip(probs, outs, blocks):
        single = BTE_RTA(pr['geo'], pr['mat'], pr['bcs'],
                         effective_thermal_conductivity=pr['effective_thermal_conductivity'],
                         bte_max_iter=4, config=SweepConfig(mc=mc, block=256))
        for o in (out, blk):
            assert o.aux['n_iter'] == single.aux['n_iter']
            assert o.kappa_eff == single.kappa_eff
            np.testing.assert_array_equal(o.variables['Temperature_BTE']['data'],
                                          single.variables['Temperature_BTE']['data'])
            np.testing.assert_array_equal(o.variables['Flux_BTE']['data'], single.variables['Flux_BTE']['data'])
            np.testing.assert_array_equal(o.aux['suppression'], single.aux['suppression'])


def test_batch_different_iteration_counts():
    """Problems stop at their own iteration; the early ones stay frozen."""
    from openbte_amd.bte import BTE_RTA
    from openbte_amd.batch import BTE_RTA_batch
    probs = [_problem(21, 400), _problem(22, 400)]
    outs = BTE_RTA_batch(probs, bte_max_iter=30, bte_rtol=1e-3, autotune=False)
    for pr, out in zip(probs, outs):
        single = BTE_RTA(pr['geo'], pr['mat'], pr['bcs'],
                         effective_thermal_conductivity=pr['effective_thermal_conductivity'],
                         bte_max_iter=30, bte_rtol=1e-3)
        assert out.aux['n_iter'] == single.aux['n_iter']
        np.testing.assert_array_equal(out.variables['Temperature_BTE']['data'],
                                      single.variables['Temperature_BTE']['data'])


def test_batch_graph_replay_matches_eager():
    """hipGraph-replayed outer iterations give the eager results, replay after replay."""
    from openbte_amd.batch import BTEBatch
    from openbte_amd.operator import build_host_operator
    probs = [_problem(31, 500), _problem(32, 700)]
    ops = [build_host_operator(p['geo'], p['mat'], p['bcs'], p['effective_thermal_conductivity']) for p in probs]
    t0 = [np.zeros(p['geo'].n_elems) for p in probs]
    eager = BTEBatch(ops)
    graph = BTEBatch(ops)
    graph.enable_graphs(True)
    eager.reset(t0)
    graph.reset(t0)
    for it in range(6):
        eager.outer_iteration()
        graph.outer_iteration()
        for de, dg in zip(eager.devs, graph.devs):
            np.testing.assert_array_equal(de.buf['dt'][0].cpu().numpy(), dg.buf['dt'][0].cpu().numpy())
            np.testing.assert_array_equal(de.buf['kappa'].cpu().numpy(), dg.buf['kappa'].cpu().numpy())


def test_tuner_picks_a_candidate_and_memoizes():
    """The launch-shape tuner times every candidate (whole outer iterations)
    on the actual problems, returns the fastest and reuses it for the same
    problem signature; the tuned shape solves like the default."""
    from openbte_amd import tune
    from openbte_amd.batch import BTEBatch
    from openbte_amd.bte import SweepConfig
    from openbte_amd.operator import build_host_operator
    probs = [_problem(21, 400), _problem(22, 700)]
    ops = [build_host_operator(p['geo'], p['mat'], p['bcs'], p['effective_thermal_conductivity']) for p in probs]
    cands = ((4, 128), (8, 256), (16, 256))
    print('# tune batch', flush=True)
    cfg = tune.tune(ops, candidates=cands, reps=2)
    res = tune.last_results(ops)
    assert (cfg.mc, cfg.block) in cands
    assert len(res) == len(cands) and all(r['iter_s'] > 0 for r in res)
    assert min(res, key=lambda r: r['iter_s'])['mc'] == cfg.mc
    again = tune.tune(ops, candidates=cands)
    assert (again.mc, again.block) == (cfg.mc, cfg.block) and tune.last_results(ops) is res
    print('# tune single', flush=True)
    one = tune.tune(ops[:1], candidates=cands, reps=2)
    assert (one.mc, one.block) in cands
    t0s = [np.zeros(o.n) for o in ops]
    ks = []
    for c in (cfg, SweepConfig()):
        b = BTEBatch(ops, cfg=SweepConfig(mc=c.mc, block=c.block))
        b.reset(t0s)
        for _ in range(3):
            b.outer_iteration()
        ks.append([float(d.buf['kappa'].sum().item()) for d in b.devs])
        del b
    np.testing.assert_allclose(ks[0], ks[1], rtol=1e-9)


def test_batch_default_tuned_graphs_match_oracle():
    """The public entry with its defaults -- launch shape from the tuner,
    hipGraph-replayed outer iterations -- against the oracle (kappa 1e-6,
    T 1e-8 K, same iteration counts)."""
    from oracle import bte_oracle as bo, fourier_oracle as fo
    from openbte_amd.batch import BTE_RTA_batch
    probs = [_problem(41, 600), _problem(42, 1200), _problem(43, 300)]
    for pr in probs:
        pr['fourier'] = fo.fourier(pr['geo'], pr['mat'].thermal_conductivity, pr['bcs'],
                                   pr['effective_thermal_conductivity'])
    outs = BTE_RTA_batch(probs, bte_max_iter=10, bte_rtol=1e-5)
    for pr, out in zip(probs, outs):
        ref = bo.bte_rta(pr['geo'], pr['mat'], pr['bcs'], fourier=pr['fourier'],
                         effective_thermal_conductivity=pr['effective_thermal_conductivity'], bte_max_iter=10,
                         bte_rtol=1e-5)
        print('tuned batch', out.kappa_eff, ref.kappa_eff, out.aux['n_iter'], flush=True)
        assert out.aux['n_iter'] == ref.aux['n_iter']
        assert abs(out.kappa_eff / ref.kappa_eff - 1) < 1e-6
        assert np.abs(out.variables['Temperature_BTE']['data']
                      - ref.variables['Temperature_BTE']['data']).max() < 1e-8


@pytest.mark.parametrize('mc', [8, 4])
def test_batch_fused_extension_equals_rounds(mc):
    """The batched extension loop in one launch (obte_batch_extend: each task
    loops begin, transfer sweep and end on its own) gives the bits and the
    extension counts of the round-by-round loop, eager and graph-replayed,
    with workgroup and wave tasks in the batch."""
    from openbte_amd.batch import BTEBatch
    from openbte_amd.bte import SweepConfig
    from openbte_amd.operator import build_host_operator
    probs = [_problem(51, 8000), _problem(52, 300), _problem(53, 900)]
    ops = [build_host_operator(p['geo'], p['mat'], p['bcs'], p['effective_thermal_conductivity']) for p in probs]
    t0 = [np.zeros(p['geo'].n_elems) for p in probs]
    runs = []
    for fused, graphs in ((False, False), (True, False), (True, True)):
        print(f'# batch mc={mc} fused={fused} graphs={graphs}', flush=True)
        b = BTEBatch(ops, cfg=SweepConfig(mc=mc, block=256, fused_extend=fused))
        if graphs:
            b.enable_graphs(True)
        b.reset(t0)
        exts = [b.outer_iteration() for _ in range(4)]
        runs.append((exts, [d.buf['dt'][0].cpu().numpy() for d in b.devs],
                     [d.buf['kappa'].cpu().numpy() for d in b.devs], [d.buf['j'].cpu().numpy() for d in b.devs]))
        del b
    ref = runs[0]
    assert ref[0][0] > 1
    for r in runs[1:]:
        assert r[0] == ref[0]
        for k in (1, 2, 3):
            for a, e in zip(r[k], ref[k]):
                np.testing.assert_array_equal(a, e)
