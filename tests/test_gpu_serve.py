"""Continuous-batching server: every served result matches the oracle
(kappa 1e-6 relative, temperature 1e-8 K) and equals the result of the same
solve alone (same launch shape), whatever else shares the device, also for a
mix that holds a 200k-triangle cell and a ~150k-tetrahedron slab; admission
respects the HBM budget; the footprint estimate is exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems():
    from oracle import fourier_oracle as fo
    from openbte_amd.geometry import pore_cell_2d, pore_slab_3d
    from openbte_amd.material import synthetic_material_2d, synthetic_material_3d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    out = []
    for i, (n, phi, M) in enumerate([(500, 0.2, 6), (900, 0.3, 7), (400, 0.1, 5)]):
        geo = pore_cell_2d(10.0, phi, n, seed=10 + i)
        mat = synthetic_material_2d(M, 8, 1.0, 1.0e4)
        eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
        out.append((geo, mat, bcs, eff))
    geo = pore_slab_3d(10.0, 3.0, 0.2, 900, seed=2, nz=3)
    mat = synthetic_material_3d(5, 2, 4, 1.0, 1e3)
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0 / 3.0)
    out.append((geo, mat, bcs, eff))
    fs = [fo.fourier(g, m.thermal_conductivity, b, e) for g, m, b, e in out]
    return out, fs


@pytest.fixture(scope='module')
def alone():
    from openbte_amd.bte import BTE_RTA, SweepConfig
    probs, fs = _problems()
    cfg = SweepConfig(mc=8, block=128)
    res = [BTE_RTA(g, m, b, fourier=f, effective_thermal_conductivity=e, bte_max_iter=12, bte_rtol=1e-6,
                   config=SweepConfig(**vars(cfg)))
           for (g, m, b, e), f in zip(probs, fs)]
    return probs, fs, cfg, res


def _same(a, b):
    assert a.aux['n_iter'] == b.aux['n_iter']
    assert abs(a.kappa_eff - b.kappa_eff) <= 1e-13 * abs(b.kappa_eff)
    dT = np.abs(a.variables['Temperature_BTE']['data'] - b.variables['Temperature_BTE']['data']).max()
    assert dT <= 1e-13
    dJ = np.abs(a.variables['Flux_BTE']['data'] - b.variables['Flux_BTE']['data']).max()
    assert dJ <= 1e-12 * np.abs(b.variables['Flux_BTE']['data']).max()


def test_served_results_equal_alone(alone):
    from openbte_amd.bte import SweepConfig
    from openbte_amd.serve import BTEServer, Request
    probs, fs, cfg, ref = alone
    srv = BTEServer(cfg=SweepConfig(**vars(cfg)), max_slots=8)
    reqs = [Request(g, m, b, e, fourier=f, bte_max_iter=12, bte_rtol=1e-6) for (g, m, b, e), f in zip(probs, fs)]
    # staggered arrivals: two at once, the others while those run
    srv.submit(reqs[0])
    srv.submit(reqs[1])
    srv.wait_prepared()
    srv.step()
    assert len(srv.live) == 2
    srv.step()
    srv.submit(reqs[2])
    srv.wait_prepared()
    srv.step()
    srv.submit(reqs[3])
    srv.wait_prepared()
    srv.run_until_idle()
    srv.close()
    assert srv.stats.peak_live >= 3
    for r, a in zip(reqs, ref):
        assert r.error is None
        print('served', r.id, r.result.kappa_eff, 'alone', a.kappa_eff, 'latency', r.latency)
        _same(r.result, a)


@pytest.fixture(scope='module')
def oracle_results(alone):
    from oracle import bte_oracle as bo
    probs, fs, cfg, _ = alone
    return [bo.bte_rta(g, m, b, fourier=f, effective_thermal_conductivity=e, bte_max_iter=12, bte_rtol=1e-6,
                       n_workers=4) for (g, m, b, e), f in zip(probs, fs)]


def test_served_results_match_oracle(alone, oracle_results):
    """Every request of a staggered served mix (three 2D pore cells and a
    tetrahedral slab) against the oracle's SuperLU solve of the same inputs
    with the same outer-iteration count."""
    from openbte_amd.bte import SweepConfig
    from openbte_amd.serve import BTEServer, Request
    probs, fs, cfg, _ = alone
    srv = BTEServer(cfg=SweepConfig(**vars(cfg)), max_slots=8)
    reqs = [Request(g, m, b, e, fourier=f, bte_max_iter=12, bte_rtol=1e-6) for (g, m, b, e), f in zip(probs, fs)]
    srv.submit(reqs[3])
    srv.submit(reqs[0])
    srv.wait_prepared()
    srv.step()
    srv.submit(reqs[1])
    srv.submit(reqs[2])
    srv.run_until_idle()
    srv.close()
    for r, o in zip(reqs, oracle_results):
        assert r.error is None
        assert r.result.aux['n_iter'] == o.aux['n_iter']
        rel = abs(r.result.kappa_eff / o.kappa_eff - 1)
        dT = np.abs(r.result.variables['Temperature_BTE']['data'] - o.variables['Temperature_BTE']['data']).max()
        Jo = o.variables['Flux_BTE']['data']
        dJ = np.abs(r.result.variables['Flux_BTE']['data'] - Jo).max() / np.abs(Jo).max()
        print('served vs oracle', r.id, r.result.kappa_eff, o.kappa_eff, rel, dT, dJ)
        assert rel < 1e-6 and dT < 1e-8 and dJ < 1e-8


def test_admission_by_footprint(alone):
    from openbte_amd.bte import SweepConfig, device_bytes
    from openbte_amd.operator import build_host_operator, build_schedule
    from openbte_amd.serve import BTEServer, Request
    probs, fs, cfg, ref = alone
    sizes = []
    for g, m, b, e in probs:
        op = build_host_operator(g, m, b, e)
        sizes.append(device_bytes(op, build_schedule(op), cfg))
    budget = max(sizes) + min(sizes) // 2  # never two of the larger problems at once
    srv = BTEServer(cfg=SweepConfig(**vars(cfg)), hbm_budget=budget, max_slots=8)
    reqs = [srv.submit(Request(g, m, b, e, fourier=f, bte_max_iter=12, bte_rtol=1e-6))
            for (g, m, b, e), f in zip(probs, fs)]
    srv.run_until_idle()
    srv.close()
    assert srv.stats.peak_bytes <= budget
    for r, a in zip(reqs, ref):
        _same(r.result, a)
    # a request larger than the whole budget fails alone, the server goes on
    small = BTEServer(cfg=SweepConfig(**vars(cfg)), hbm_budget=min(sizes) - 1, max_slots=2)
    g, m, b, e = probs[0]
    r = small.submit(Request(g, m, b, e, fourier=fs[0]))
    small.run_until_idle()
    small.close()
    assert isinstance(r.error, MemoryError) and r.result is None


def test_footprint_estimate_is_exact(alone):
    from openbte_amd.bte import BTEDevice, SweepConfig, device_bytes
    from openbte_amd.operator import build_host_operator, build_schedule
    probs, fs, cfg, ref = alone
    for g, m, b, e in probs:
        op = build_host_operator(g, m, b, e)
        sc = build_schedule(op)
        d = BTEDevice(op, sched=sc, cfg=SweepConfig(**vars(cfg)))
        assert device_bytes(op, sc, cfg) == d.nbytes()


def test_full_size_mix_served_equals_alone():
    """The serving workload's largest requests -- a 200k-triangle pore cell
    (20 MFP x 48 directions) and a ~150k-tetrahedron porous slab (20 MFP x
    64 directions) -- served together with two small requests arriving while
    they run: each result equals its solve alone bit for bit (the kernels treat
    problems independently and reduce in a fixed order)."""
    import torch
    from openbte_amd.bte import BTE_RTA, SweepConfig
    from openbte_amd.fourier import Fourier
    from openbte_amd.geometry import pore_cell_2d, pore_slab_3d
    from openbte_amd.material import synthetic_material_2d, synthetic_material_3d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    from openbte_amd.serve import BTEServer, Request
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff2 = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    L3 = float(np.sqrt(150_000 * 0.45 ** 3 / 0.8 / 4.0) ** (2 / 3) * 4.0 ** (1 / 3))
    eff3 = EffectiveThermalConductivity(contact='Periodic_x', normalization=-4.0 / L3)
    probs = [(pore_cell_2d(100.0, 0.25, 200_000, seed=21), synthetic_material_2d(20, 48, 1.0, 1.0e4, seed=7), eff2),
             (pore_slab_3d(L3, L3 / 4.0, 0.2, 150_000, seed=22), synthetic_material_3d(20, 8, 8, 1.0, 1.0e4, seed=7),
              eff3),
             (pore_cell_2d(10.0, 0.2, 1500, seed=23), synthetic_material_2d(30, 48, 1.0, 1.0e4, seed=7), eff2),
             (pore_slab_3d(10.0, 3.0, 0.2, 6000, seed=24), synthetic_material_3d(24, 8, 8, 1.0, 1.0e4, seed=7),
              EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0 / 3.0))]
    print('# full-size mix:', [g.n_elems for g, _, _ in probs], flush=True)
    assert probs[0][0].n_elems > 190_000 and probs[1][0].n_elems > 140_000
    fs = [Fourier(g, m.thermal_conductivity, bcs, e, verbose=False) for g, m, e in probs]
    cfg = SweepConfig()
    alone = [BTE_RTA(g, m, bcs, fourier=f, effective_thermal_conductivity=e, config=SweepConfig(**vars(cfg)),
                     verbose=False) for (g, m, e), f in zip(probs, fs)]
    torch.cuda.empty_cache()
    srv = BTEServer(cfg=SweepConfig(**vars(cfg)), max_slots=8)
    reqs = [Request(g, m, bcs, e, fourier=f) for (g, m, e), f in zip(probs, fs)]
    srv.submit(reqs[0])
    srv.submit(reqs[1])
    srv.wait_prepared()
    srv.step()
    srv.step()
    srv.submit(reqs[2])
    srv.submit(reqs[3])
    srv.run_until_idle()
    srv.close()
    assert srv.stats.peak_live == 4
    for r, a in zip(reqs, alone):
        assert r.error is None
        print('full-size served', r.id, r.result.kappa_eff, 'alone', a.kappa_eff, r.result.aux['n_iter'], flush=True)
        assert r.result.aux['n_iter'] == a.aux['n_iter']
        assert r.result.kappa_eff == a.kappa_eff
        assert np.array_equal(r.result.variables['Temperature_BTE']['data'], a.variables['Temperature_BTE']['data'])
        assert np.array_equal(r.result.variables['Flux_BTE']['data'], a.variables['Flux_BTE']['data'])


def test_cancel_timeout_and_failure_leave_others_unchanged(alone):
    """Host-forced endings: one request cancelled while it runs, one whose
    deadline passes while it runs (timeout_s set to 0 after admission), one
    whose preparation fails (unknown contact name); the other two finish with
    the results they have alone, and every slot and byte is released."""
    from openbte_amd.bte import SweepConfig
    from openbte_amd.objects import EffectiveThermalConductivity
    from openbte_amd.serve import BTEServer, Request, RequestCancelled, RequestTimeout
    probs, fs, cfg, ref = alone
    srv = BTEServer(cfg=SweepConfig(**vars(cfg)), max_slots=8)
    reqs = [Request(g, m, b, e, fourier=f, bte_max_iter=12, bte_rtol=1e-6) for (g, m, b, e), f in zip(probs, fs)]
    g, m, b, _ = probs[0]
    bad = Request(g, m, b, EffectiveThermalConductivity(contact='NoSuchRegion', normalization=-1.0), fourier=fs[0])
    for r in reqs + [bad]:
        srv.submit(r)
    srv.wait_prepared()
    srv.step()
    srv.step()
    assert len(srv.live) == 4
    srv.cancel(reqs[1])
    reqs[2].timeout_s = 0.0
    srv.run_until_idle()
    srv.close()
    assert isinstance(reqs[1].error, RequestCancelled) and reqs[1].result is None
    assert isinstance(reqs[2].error, RequestTimeout) and reqs[2].result is None
    assert bad.error is not None and bad.result is None
    for k in (0, 3):
        assert reqs[k].error is None
        _same(reqs[k].result, ref[k])
    assert srv.used_bytes == 0 and len(srv.free_slots) == 8
    assert srv.stats.failed == 3 and srv.stats.completed == 5
