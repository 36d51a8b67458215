"""Benchmark of the steady-state mode-resolved BTE solve on one MI355X.

Workload (default ``--config plain200k``): periodic square cell L = 100 nm with
one centred circular pore (porosity 0.25, diffuse walls), ~200,000-cell
seeded jittered-Delaunay triangle mesh, 60 MFP bins (1 nm - 100 um) x 96
directions = 5,760 upwind systems per outer iteration, FP64, DeltaT = 1 K
along x.  Synthetic data (seeded mesh and material).

A step is one outer iteration of the solve (all 5,760 systems solved to
direct-solver accuracy + mode sums).  Warmup steps run on a throwaway solve;
the K timed steps are outer iterations 1..K of a fresh solve started from the
Fourier guess with an empty Krylov basis, so the timed region contains the
expensive first iterations.  ``value`` = n * M * L * K / seconds.

Also reported:
  roofline      sweep kernel (the dominant kernel): algorithmic bytes per
                launch (28 B per cell*mode update for triangles) / mean launch
                duration from HIP events on the launch stream, vs 8 TB/s.
  cpu_baseline  the FP64 SuperLU restatement (oracle) timed on this host's
                cores on a bounded sample of the same systems (steady-state
                cached-LU solves; factorisation reported separately).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BATCH_CONFIGS = {
    # name: (geometries, min cells, max cells, n_mfp, n_dirs, mfp_max nm, seed)
    'batch64': (64, 1000, 80000, 30, 48, 1.0e4, 2026),
}

SERVE_CONFIGS = {
    # name: (requests, 2D cells range, 3D tets range, MFP bins range, 2D dirs, 3D (theta, phi), 3D share, seed)
    'serve': (24, (1000, 200_000), (5000, 150_000), (20, 60), 48, (8, 8), 0.4, 7),
}

INVERSE_CONFIGS = {
    # name: (pixels per side, L nm, n_mfp, n_dirs, mfp_max nm, smoothing (pixels), seed)
    'inverse40': (40, 100.0, 30, 48, 1.0e4, 3.0, 11),
}

CONFIGS = {
    # name: (L nm, porosity, cells, n_mfp, n_dirs, mfp_max nm)
    'plain2k': (10.0, 0.25, 2000, 30, 48, 1.0e4),
    'plain200k': (100.0, 0.25, 200_000, 60, 96, 1.0e5),
}
CONFIGS3D = {
    # name: (tetrahedra, porosity, n_mfp, (n_theta, n_phi), mfp_max nm): a periodic
    # porous slab (cylindrical pore, diffuse pore wall and slab faces)
    'tet150k': (150_000, 0.2, 40, (8, 8), 1.0e4),
}
HBM_PEAK = 8.0e12
BASELINE_METRIC = 'cell\u00b7mode updates/s + p50/p99 solve latency vs plain GPU, kappa rel err \u22641e-6'


def make_case(name, seed=0):
    from openbte_amd.geometry import pore_cell_2d, pore_slab_3d
    from openbte_amd.material import synthetic_material_2d, synthetic_material_3d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    if name in CONFIGS3D:
        n, phi, M, (nth, nph), mfp_max = CONFIGS3D[name]
        L = float(np.sqrt(n * 0.45 ** 3 / (1 - phi) / 4.0) ** (2 / 3) * 4.0 ** (1 / 3))
        H = L / 4.0
        geo = pore_slab_3d(L, H, phi, n, seed=seed)
        mat = synthetic_material_3d(M, nth, nph, 1.0, mfp_max, seed=seed)
        bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
        eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0 / H)
        return geo, mat, bcs, eff
    L, por, n, M, nphi, mfp_max = CONFIGS[name]
    geo = pore_cell_2d(L, por, n, seed=seed)
    mat = synthetic_material_2d(M, nphi, 1.0, mfp_max, seed=seed)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-L / L)
    return geo, mat, bcs, eff


def host_cores():
    """Host cores the CPU legs may use: this process's affinity set, capped at
    16 (the host share of one GPU on a shared multi-GPU node)."""
    aff = len(os.sched_getaffinity(0))
    return min(aff, 16), aff


def cpu_baseline(geo, mat, bcs, eff, fourier_T, budget_s=30.0):
    """Time the oracle's SuperLU path on a bounded, stratified sample of the
    workload's systems: every MFP bin once, at directions spread over the
    whole set (bin m at direction (37 m) mod L), dealt round-robin to the
    workers; each worker factorises its systems, then repeats cached solves.
    Updates/s = n * systems solved / wall seconds (factor and solve timed
    separately)."""
    import multiprocessing as mp
    import scipy.sparse.linalg as spla
    from oracle import bte_oracle as bo
    ops = bo.assemble(geo, mat, bcs, eff)
    cores, aff = host_cores()
    n, M, L = geo.n_elems, ops.n_mfp, ops.n_dirs
    TB = fourier_T[ops.i_diffuse].copy()
    systems = [((37 * m) % L, m) for m in range(M)]
    while len(systems) < cores:  # every worker gets at least one system
        k = len(systems)
        systems.append(((37 * k + 11) % L, k % M))
    per_worker = [systems[w::cores] for w in range(cores)]
    # probe one solve to size the cached-solve repetitions
    lu = spla.splu(bo.system_matrix(ops, systems[0][0], systems[0][1]))
    t0 = time.perf_counter()
    lu.solve(bo.rhs(ops, systems[0][0], systems[0][1], fourier_T, TB))
    t_sol = time.perf_counter() - t0
    k_max = max(len(s) for s in per_worker)
    reps = int(max(1, (0.5 * budget_s) // max(k_max * t_sol, 1e-4)))
    ctx = mp.get_context('fork')
    q = ctx.Queue()

    def work(w):
        mine = per_worker[w]
        t0 = time.perf_counter()
        lus = [spla.splu(bo.system_matrix(ops, l, m)) for l, m in mine]
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(reps):
            for (l, m), f in zip(mine, lus):
                f.solve(bo.rhs(ops, l, m, fourier_T, TB))
        q.put((w, tf, time.perf_counter() - t0))

    procs = [ctx.Process(target=work, args=(w,)) for w in range(cores)]
    wall0 = time.perf_counter()
    for p in procs:
        p.start()
    res = [q.get() for _ in procs]
    for p in procs:
        p.join()
    wall = time.perf_counter() - wall0
    tf = max(r[1] for r in res)
    ts = max(r[2] for r in res)
    n_sys = len(systems)
    steady = n * n_sys * reps / ts
    first = n * n_sys / (tf + ts / reps)
    return {
        'value': steady, 'unit': 'cell*mode updates/s', 'cores': cores, 'kind': 'port',
        'sample': (f'{n_sys} (direction, MFP bin) systems of this workload covering all {M} MFP bins at '
                   f'{len(set(l for l, _ in systems))} directions, dealt to {cores} forked workers '
                   f'({aff} cores in the affinity set; at most 16 used, the host share of one GPU); '
                   f'SuperLU factor, then {reps} cached solves each; value = steady-state (cached LU) rate; '
                   f'first-iteration rate (factor + solve) {first:.3e}; wall {wall:.1f}s'),
        'first_iteration_value': first, 'affinity_cores': aff,
    }


def c1_oracle(workers):
    """The CPU-runnable workload (2k-cell pore cell, 30 MFP x 48 directions)
    solved by the oracle from its own Fourier guess until kappa changes by
    less than 1e-6 relative between outer iterations; wall-clock timed."""
    from oracle import bte_oracle as bo
    from oracle import fourier_oracle as fo
    geo, mat, bcs, eff = make_case('plain2k')
    t0 = time.perf_counter()
    f = fo.fourier(geo, mat.thermal_conductivity, bcs, eff)
    ops = bo.assemble(geo, mat, bcs, eff)
    T_F = np.asarray(f.variables['Temperature_Fourier']['data'], np.float64)
    DeltaT, TB = T_F.copy(), T_F[ops.i_diffuse].copy()
    kappas = []
    with bo.OuterLoop(ops, workers) as loop:
        while len(kappas) < 400:
            DeltaT, TB, J, K, S = loop.step(DeltaT, TB)
            kappas.append(float(K.sum()))
            if len(kappas) > 1 and abs(kappas[-1] - kappas[-2]) <= 1e-6 * abs(kappas[-1]):
                break
    wall = time.perf_counter() - t0
    n, M, L = geo.n_elems, ops.n_mfp, ops.n_dirs
    return {'fourier_T': T_F, 'delta_t': DeltaT, 'kappa': kappas[-1], 'n_iter': len(kappas), 'seconds': wall,
            'updates_per_s': n * M * L * len(kappas) / wall, 'workers': workers, 'n': n}


def c1_gpu_check(c1):
    """The same 2k-cell solve on the GPU, from the oracle's Fourier guess,
    for the oracle's iteration count: kappa and temperature errors."""
    import torch
    from openbte_amd.bte import BTE_RTA
    from openbte_amd.objects import SolverResults
    geo, mat, bcs, eff = make_case('plain2k')
    f = SolverResults(variables={'Temperature_Fourier': {'data': c1['fourier_T'], 'units': 'K'}})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = BTE_RTA(geo, mat, bcs, fourier=f, effective_thermal_conductivity=eff, bte_max_iter=c1['n_iter'],
                bte_rtol=0.0, verbose=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {'kappa_gpu': r.kappa_eff, 'kappa_oracle': c1['kappa'],
            'kappa_rel_err': abs(r.kappa_eff / c1['kappa'] - 1),
            'temperature_max_abs_err_K': float(np.abs(r.variables['Temperature_BTE']['data'] - c1['delta_t']).max()),
            'outer_iterations': c1['n_iter'], 'gpu_solve_s': wall, 'oracle_solve_s': c1['seconds'],
            'oracle_workers': c1['workers'], 'oracle_updates_per_s': c1['updates_per_s']}


def make_batch(name):
    from openbte_amd.geometry import random_pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    G, nmin, nmax, M, nphi, mfp_max, seed = BATCH_CONFIGS[name]
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), G)).astype(int)
    mat = synthetic_material_2d(M, nphi, 1.0, mfp_max, seed=seed)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    geos = [random_pore_cell_2d(seed + 1 + i, int(n)) for i, n in enumerate(ns)]
    return geos, mat, bcs, eff


def measure_batch(args):
    """Batched form vs the plain one-solve-at-a-time form on the same problems."""
    import torch
    from openbte_amd.batch import BTEBatch
    from openbte_amd.bte import BTEDevice, SweepConfig
    from openbte_amd.fourier import Fourier
    from openbte_amd.operator import DEFAULT_MAX_WIDTH, batch_max_width, build_host_operator, build_schedule
    geos, mat, bcs, eff = make_batch(args.config)
    M, L = mat.grid[0], mat.grid[1]
    ncells = np.array([g.n_elems for g in geos])
    print(f'# {args.config}: {len(geos)} geometries, {ncells.sum()} cells ({ncells.min()}-{ncells.max()})',
          file=sys.stderr, flush=True)
    t0s = []
    for g in geos:
        f = Fourier(g, mat.thermal_conductivity, bcs, eff, verbose=False)
        t0s.append(f.variables['Temperature_Fourier']['data'])
    ops = [build_host_operator(g, mat, bcs, eff) for g in geos]
    tuned = None
    if args.mc or args.block:
        cfg = SweepConfig(mc=args.mc or 8, block=args.block or 256, window=not args.no_window)
    else:
        from openbte_amd import tune
        cfg = tune.tune(ops)
        cfg.window = not args.no_window
        tuned = tune.last_results(ops)
        print(f'# autotuned mc={cfg.mc} block={cfg.block}', file=sys.stderr, flush=True)
    # the batched form's schedules (levels of at most one batch-kernel pass);
    # the plain form is timed with these and with the single-solve default's
    width = batch_max_width(cfg.mc, cfg.block)
    scheds = [build_schedule(o, max_width=width) for o in ops]
    scheds_by_width = {width: scheds, DEFAULT_MAX_WIDTH: [build_schedule(o) for o in ops]}
    cfg.fused_extend = not args.no_fused
    K = args.steps
    updates = float(ncells.sum()) * M * L * K

    batch = BTEBatch(ops, cfg=SweepConfig(**vars(cfg)), scheds=scheds)
    if not args.no_graphs:
        batch.enable_graphs(True)
    batch.reset(t0s)
    for _ in range(args.warmup):
        batch.outer_iteration()
    batch.reset(t0s)
    torch.cuda.synchronize()
    t = time.perf_counter()
    exts = [batch.outer_iteration() for _ in range(K)]
    torch.cuda.synchronize()
    dt_batch = time.perf_counter() - t
    kappas = [float(d.buf['kappa'].sum().item()) for d in batch.devs]
    del batch
    torch.cuda.empty_cache()

    # plain form: the same K outer iterations, one geometry at a time, at the
    # plain path's default shape and at the batch's shape; the faster counts
    def plain(pcfg, width=DEFAULT_MAX_WIDTH):
        dt, ks = 0.0, []
        for op, sc, t0 in zip(ops, scheds_by_width[width], t0s):
            dev = BTEDevice(op, sched=sc, cfg=SweepConfig(**vars(pcfg)))
            dev.reset(t0)
            dev.outer_iteration()
            dev.reset(t0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(K):
                dev.outer_iteration()
            torch.cuda.synchronize()
            dt += time.perf_counter() - t
            ks.append(float(dev.buf['kappa'].sum().item()))
            del dev
        torch.cuda.empty_cache()
        return dt, ks
    # the plain form at the plain path's default shape, at mc 4 (faster for
    # small meshes alone) and at the batch's shape, each with the single-solve
    # default schedules and with the batched form's; the fastest counts
    plain_runs = {}
    for w in scheds_by_width:
        plain_runs[(8, 256, w)] = plain(SweepConfig(window=not args.no_window, fused_extend=not args.no_fused), w)
        plain_runs[(4, 256, w)] = plain(SweepConfig(mc=4, window=not args.no_window,
                                                    fused_extend=not args.no_fused), w)
        if (cfg.mc, cfg.block, w) not in plain_runs:
            plain_runs[(cfg.mc, cfg.block, w)] = plain(cfg, w)
    best_plain = min(plain_runs, key=lambda k: plain_runs[k][0])
    dt_plain = plain_runs[best_plain][0]
    plain_kappa = plain_runs[(cfg.mc, cfg.block, DEFAULT_MAX_WIDTH)][1]
    same = max(abs(a - b) / abs(b) for a, b in zip(kappas, plain_kappa))
    value = updates / dt_batch
    out = {
        'metric': 'cell*mode updates/s (batched outer iterations, all systems to direct-solver accuracy)',
        'value': value, 'unit': 'cell*mode updates/s', 'n_gpus': 1, 'steps': K, 'warmup': args.warmup,
        'ms_per_step': dt_batch / K * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'synthetic (64 seeded multi-pore cells, seeded lognormal Si-like material)',
        'config': {'workload': f'{args.config}: {len(geos)} periodic multi-pore cells, {int(ncells.sum())} cells '
                               f'({int(ncells.min())}-{int(ncells.max())}), {M} MFP x {L} directions',
                   'mc': cfg.mc, 'block': cfg.block, 'hipgraph': not args.no_graphs, 'fused_extend': cfg.fused_extend,
                   'extensions_per_step': exts,
                   'autotune': tuned,
                   'plain_value': updates / dt_plain, 'plain_ms_per_step': dt_plain / K * 1e3,
                   'plain_shape': list(best_plain),
                   'plain_values_by_shape': {f'mc{k[0]}_block{k[1]}_width{k[2]}': updates / v[0]
                                             for k, v in plain_runs.items()},
                   'schedule_max_width': width,
                   'batched_over_plain': dt_plain / dt_batch,
                   'max_rel_kappa_diff_batched_vs_plain': same},
    }
    return out


def run_batch(args):
    print(json.dumps(measure_batch(args)))
    return 0


def make_requests(name):
    """Seeded request mix: 2D pore cells and 3D porous slabs (tetrahedra)."""
    from openbte_amd.geometry import pore_cell_2d, pore_slab_3d
    from openbte_amd.material import synthetic_material_2d, synthetic_material_3d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    N, r2, r3, rm, nphi, (nth, nph), share3, seed = SERVE_CONFIGS[name]
    rng = np.random.default_rng(seed)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    out = []
    for i in range(N):
        M = int(rng.integers(rm[0], rm[1] + 1))
        phi = float(rng.uniform(0.05, 0.35))
        if rng.random() < share3:
            n = int(np.exp(rng.uniform(np.log(r3[0]), np.log(r3[1]))))
            L = float(np.sqrt(n * 0.45 ** 3 / (1 - phi) / 4.0) ** (2 / 3) * 4.0 ** (1 / 3))
            H = L / 4.0
            geo = pore_slab_3d(L, H, phi, n, seed=seed + i)
            mat = synthetic_material_3d(M, nth, nph, 1.0, 1.0e4, seed=seed)
            eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0 / H)
        else:
            n = int(np.exp(rng.uniform(np.log(r2[0]), np.log(r2[1]))))
            L = float(np.sqrt(n * 0.274 ** 2 / (2.0 * (1.0 - phi))))
            geo = pore_cell_2d(L, phi, n, seed=seed + i)
            mat = synthetic_material_2d(M, nphi, 1.0, 1.0e4, seed=seed)
            eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
        out.append((geo, mat, bcs, eff))
    return out


def measure_serve(args):
    """Continuous-batching server vs the plain one-solve-at-a-time GPU form
    under the same open-loop Poisson request stream (lambda = 70% of the
    server's measured saturated capacity)."""
    import torch
    from openbte_amd.bte import BTE_RTA, SweepConfig
    from openbte_amd.fourier import Fourier
    from openbte_amd.serve import BTEServer, Request

    t0 = time.perf_counter()
    probs = make_requests(args.config)
    N = len(probs)
    sizes = [(g.n_elems, g.dim, m.t_coeff.shape[0], m.t_coeff.shape[1]) for g, m, _, _ in probs]
    print(f'# {args.config}: {N} requests generated in {time.perf_counter() - t0:.1f}s; '
          f'cells {min(s[0] for s in sizes)}-{max(s[0] for s in sizes)}, '
          f'{sum(s[1] == 3 for s in sizes)} 3D', file=sys.stderr, flush=True)
    cfg = SweepConfig()
    if args.mc:
        cfg.mc = args.mc
    if args.block:
        cfg.block = args.block
    rtol, maxit = 1e-4, 100

    # warm-up: one small solve of each kind (kernel loading, allocator)
    for g, m, b, e in sorted(probs, key=lambda p: p[0].n_elems)[:1] + [p for p in probs if p[0].dim == 3][:1]:
        f = Fourier(g, m.thermal_conductivity, b, e, verbose=False)
        BTE_RTA(g, m, b, fourier=f, effective_thermal_conductivity=e, bte_max_iter=2,
                config=SweepConfig(**vars(cfg)), verbose=False)
    torch.cuda.synchronize()

    # plain form: each request alone (Fourier guess + BTE solve), one after the other
    service, plain_res, updates = [], [], []
    for g, m, b, e in probs:
        torch.cuda.synchronize()
        ts = time.perf_counter()
        f = Fourier(g, m.thermal_conductivity, b, e, verbose=False)
        r = BTE_RTA(g, m, b, fourier=f, effective_thermal_conductivity=e, bte_max_iter=maxit, bte_rtol=rtol,
                    config=SweepConfig(**vars(cfg)), verbose=False)
        torch.cuda.synchronize()
        service.append(time.perf_counter() - ts)
        plain_res.append(r)
        M, L = m.t_coeff.shape
        updates.append(g.n_elems * M * L * r.aux['n_iter'])
    service = np.array(service)
    total_updates = float(sum(updates))
    print(f'# plain: {service.sum():.2f}s for {N} requests, {total_updates / service.sum():.3e} updates/s',
          file=sys.stderr, flush=True)

    def serve(arrivals):
        srv = BTEServer(cfg=SweepConfig(**vars(cfg)), max_slots=64, prep_workers=getattr(args, 'prep_workers', 4))
        reqs = [Request(g, m, b, e, fourier='auto', bte_max_iter=maxit, bte_rtol=rtol) for g, m, b, e in probs]
        order = np.argsort(arrivals, kind='stable')
        start = time.perf_counter()
        k = 0
        while k < N or srv.pending():
            now = time.perf_counter() - start
            while k < N and arrivals[order[k]] <= now:
                srv.submit(reqs[order[k]])
                k += 1
            got = srv.step()
            if not srv.live and not got and k < N and not srv.preparing:
                time.sleep(min(0.002, max(0.0, arrivals[order[k]] - (time.perf_counter() - start))))
        makespan = time.perf_counter() - start
        srv.close()
        return reqs, makespan, srv.stats

    # saturated capacity: every request at t = 0
    reqs_sat, span_sat, st_sat = serve(np.zeros(N))
    capacity = N / span_sat
    print(f'# served saturated: {span_sat:.2f}s, capacity {capacity:.2f} req/s, peak live {st_sat.peak_live}',
          file=sys.stderr, flush=True)
    # open loop at 70% of that capacity
    lam = 0.7 * capacity
    arrivals = np.cumsum(np.random.default_rng(123).exponential(1.0 / lam, N))
    reqs, span, st = serve(arrivals)
    lat = np.array([r.latency for r in reqs])
    prep = np.array([r.t_ready - r.t_submit for r in reqs])     # host operator, schedule, Fourier guess
    queue = np.array([r.t_admit - r.t_ready for r in reqs])     # prepared, waiting for a slot / HBM
    # the plain form under the same arrivals: a FIFO queue in front of one solver
    end, plat = 0.0, []
    for a, s_ in zip(arrivals, service):
        end = max(end, a) + s_
        plat.append(end - a)
    plat = np.array(plat)
    kdiff = max(abs(r.result.kappa_eff / p.kappa_eff - 1) for r, p in zip(reqs, plain_res))
    print(f'# served open loop: lambda {lam:.2f}/s, p50 {np.percentile(lat, 50):.3f}s p99 '
          f'{np.percentile(lat, 99):.3f}s; plain p50 {np.percentile(plat, 50):.3f}s p99 '
          f'{np.percentile(plat, 99):.3f}s; max kappa diff {kdiff:.2e}', file=sys.stderr, flush=True)
    value = total_updates / span_sat
    out = {
        'metric': 'cell*mode updates/s (served, saturated) + p50/p99 request latency at 70% load',
        'value': value, 'unit': 'cell*mode updates/s', 'n_gpus': 1, 'steps': N, 'warmup': 2,
        'ms_per_step': span_sat / N * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'synthetic (seeded 2D pore cells and 3D porous slabs, seeded Si-like material)',
        'config': {'workload': f'{args.config}: {N} requests, 2D {SERVE_CONFIGS[args.config][1]} cells x 48 '
                               f'directions, 3D {SERVE_CONFIGS[args.config][2]} tets x 64 directions, '
                               f'{SERVE_CONFIGS[args.config][3]} MFP bins, bte_rtol {rtol}',
                   'mc': cfg.mc, 'block': cfg.block,
                   'plain_value': total_updates / service.sum(),
                   'served_over_plain': service.sum() / span_sat,
                   'lambda_per_s': lam,
                   'served_p50_s': float(np.percentile(lat, 50)), 'served_p99_s': float(np.percentile(lat, 99)),
                   'plain_p50_s': float(np.percentile(plat, 50)), 'plain_p99_s': float(np.percentile(plat, 99)),
                   'served_prep_p50_s': float(np.percentile(prep, 50)),
                   'served_prep_p99_s': float(np.percentile(prep, 99)),
                   'served_queue_p50_s': float(np.percentile(queue, 50)),
                   'open_loop_updates_per_s': total_updates / span,
                   'peak_live': st_sat.peak_live, 'peak_bytes': st_sat.peak_bytes,
                   'max_rel_kappa_diff_served_vs_alone': kdiff},
    }
    return out


def run_serve(args):
    print(json.dumps(measure_serve(args)))
    return 0


def density_field(N, smooth, seed):
    """Seeded smooth density on an N x N periodic pixel grid, in [0.05, 1]."""
    z = np.random.default_rng(seed).standard_normal((N, N))
    k = np.fft.fftfreq(N)
    g = np.exp(-2 * (np.pi * smooth) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2))
    z = np.real(np.fft.ifft2(np.fft.fft2(z) * g))
    z = (z - z.mean()) / z.std()
    # smooth map into [0.05, 1]: every interface keeps some transmission
    return (0.05 + 0.95 / (1.0 + np.exp(-2.0 * z))).reshape(-1)


def run_inverse(args):
    """Differentiation workload: kappa_eff(rho) forward + adjoint gradient on
    a pixel grid (4 triangles per pixel), checked by central differences."""
    import torch
    from openbte_amd.bte import SweepConfig
    from openbte_amd.geometry import structured_tri_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    from openbte_amd.tim import KappaTIM, TIMSolver
    N, L, M, nphi, mfp_max, smooth, seed = INVERSE_CONFIGS[args.config]
    geo = structured_tri_cell_2d(L, N)
    mat = synthetic_material_2d(M, nphi, 1.0, mfp_max, seed=seed)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0})
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    rho = density_field(N, smooth, seed)
    cfg = SweepConfig()
    if args.mc:
        cfg.mc = args.mc
    if args.block:
        cfg.block = args.block
    t0 = time.perf_counter()
    solver = TIMSolver(geo, mat, bcs, eff, np.arange(geo.n_elems) // 4, cfg=cfg, rtol=1e-13)
    setup = time.perf_counter() - t0
    solver.forward(np.ones_like(rho))  # warm-up (kernels, allocator)
    solver.x_fwd = solver.x_adj = None  # timed solves start cold
    torch.cuda.synchronize()
    r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    k = KappaTIM.apply(r, solver)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    it_fwd = solver.n_iter
    t0 = time.perf_counter()
    k.backward()
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    it_adj = solver.n_iter_adj
    g = r.grad.numpy()
    # central differences (h = 1e-6) at the pixels of largest |gradient| and a random one
    picks = list(np.argsort(-np.abs(g))[:2]) + [int(np.random.default_rng(seed).integers(0, N * N))]
    h = 1e-6
    fd_err = []
    for p in picks:
        e = np.zeros_like(rho)
        e[p] = h
        fd = (solver.forward(rho + e) - solver.forward(rho - e)) / (2 * h)
        fd_err.append(abs(g[p] - fd) / np.abs(g).max())
    n = geo.n_elems
    updates = n * M * nphi * it_fwd
    value = updates / t_fwd
    print(f'# {args.config}: kappa {float(k.detach()):.6f}, forward {it_fwd} iterations {t_fwd:.2f}s, adjoint '
          f'{it_adj} iterations {t_adj:.2f}s, fd rel err {max(fd_err):.2e}', file=sys.stderr, flush=True)
    out = {
        'metric': 'cell*mode updates/s (forward TIM solve) + adjoint d kappa / d rho',
        'value': value, 'unit': 'cell*mode updates/s', 'n_gpus': 1, 'steps': it_fwd, 'warmup': 1,
        'ms_per_step': t_fwd / it_fwd * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'synthetic (seeded smoothed random density, seeded Si-like material)',
        'config': {'workload': f'{args.config}: {N}x{N} pixels, {n} triangles, {M} MFP x {nphi} directions, '
                               f'interfacial transmission t = 2 rho_a rho_b / (rho_a + rho_b)',
                   'mc': cfg.mc, 'block': cfg.block, 'kappa_eff': float(k.detach()), 'setup_s': setup,
                   'forward_s': t_fwd, 'forward_iterations': it_fwd, 'adjoint_s': t_adj,
                   'adjoint_iterations': it_adj, 'gradient_entries': int(N * N),
                   'fd_pixels': [int(p) for p in picks], 'fd_h': h,
                   'max_fd_rel_err': float(max(fd_err))},
    }
    print(json.dumps(out))
    return 0


def time_to_solution(geo, mat, bcs, eff, reps=2):
    """Wall time of a whole plain solve through the drop-in entries: Fourier
    guess, then BTE_RTA to the reference's stop rule (|dT| < 1e-4), host
    preparation and device upload included; best of ``reps`` cold solves
    (preparation caches emptied first), then one warm solve (the mesh seen
    before: schedule and Fourier preparation reused)."""
    import torch
    from openbte_amd import cache
    from openbte_amd.bte import BTE_RTA
    from openbte_amd.fourier import Fourier
    best = None
    for rep in range(reps + 1):
        warm = rep == reps
        if not warm:
            cache.clear_all()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = Fourier(geo, mat.thermal_conductivity, bcs, eff, verbose=False)
        t1 = time.perf_counter()
        r = BTE_RTA(geo, mat, bcs, fourier=f, effective_thermal_conductivity=eff, verbose=False)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        rec = {'seconds': t2 - t0, 'fourier_s': t1 - t0, 'bte_rta_s': t2 - t1,
               'bte_rta_breakdown_s': r.aux.get('timing'), 'outer_iterations': r.aux['n_iter'],
               'kappa_eff': r.kappa_eff}
        del r, f
        torch.cuda.empty_cache()
        if warm:
            best['warm_seconds'] = rec['seconds']
            best['warm'] = rec
        elif best is None or rec['seconds'] < best['seconds']:
            best = rec
    return best


def north_star_legs(args, geo, mat, bcs, eff, c1, value):
    """The rest of the headline metric beside the plain throughput: a whole
    solve's latency, the kappa error against the oracle on the converged
    2k-cell workload, and the batched and served forms against the plain
    one-solve-at-a-time form (shorter runs of the batch64 / serve configs)."""
    import torch
    torch.cuda.empty_cache()
    north = {'plain_updates_per_s': value}
    tts = time_to_solution(geo, mat, bcs, eff)
    north['plain200k_time_to_solution_s'] = tts['seconds']
    north['plain200k_time_to_solution_warm_s'] = tts['warm_seconds']
    north['plain200k_time_to_solution'] = tts
    print(f'# time to solution {tts["seconds"]:.3f}s cold, {tts["warm_seconds"]:.3f}s on a mesh seen before '
          f'({tts["outer_iterations"]} outer iterations)', file=sys.stderr, flush=True)
    if c1 is not None:
        chk = c1_gpu_check(c1)
        north['kappa_rel_err_vs_oracle'] = chk['kappa_rel_err']
        north['converged_2k_parity'] = chk
        print(f'# 2k converged: kappa rel err {chk["kappa_rel_err"]:.2e}, T err '
              f'{chk["temperature_max_abs_err_K"]:.2e} K', file=sys.stderr, flush=True)
    ba = argparse.Namespace(config='batch64', steps=5, warmup=1, mc=None, block=None,
                            no_window=args.no_window, no_graphs=False, no_fused=False)
    b = measure_batch(ba)
    north['batched_over_plain'] = b['config']['batched_over_plain']
    north['batched'] = {'value': b['value'], 'ms_per_step': b['ms_per_step'], 'steps': b['steps'],
                        'plain_value': b['config']['plain_value'], 'shape': [b['config']['mc'], b['config']['block']],
                        'max_rel_kappa_diff_batched_vs_plain': b['config']['max_rel_kappa_diff_batched_vs_plain'],
                        'workload': b['config']['workload']}
    torch.cuda.empty_cache()
    sv = measure_serve(argparse.Namespace(config='serve', mc=None, block=None, prep_workers=args.prep_workers))
    c = sv['config']
    north['served_over_plain'] = c['served_over_plain']
    north['served_p50_s'] = c['served_p50_s']
    north['served_p99_s'] = c['served_p99_s']
    north['plain_p50_s'] = c['plain_p50_s']
    north['plain_p99_s'] = c['plain_p99_s']
    north['served'] = {k: c[k] for k in ('workload', 'lambda_per_s', 'plain_value', 'open_loop_updates_per_s',
                                         'max_rel_kappa_diff_served_vs_alone', 'served_prep_p50_s', 'peak_live')}
    north['served']['value'] = sv['value']
    return north


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--config', default='plain200k',
                    choices=sorted(CONFIGS) + sorted(CONFIGS3D) + sorted(BATCH_CONFIGS) + sorted(SERVE_CONFIGS)
                    + sorted(INVERSE_CONFIGS))
    ap.add_argument('--no-graphs', action='store_true', help='batched configs: launch without hipGraphs')
    ap.add_argument('--no-window', action='store_true', help='sweeps read every upwind value from HBM')
    ap.add_argument('--no-fused', action='store_true',
                    help='batched configs: host-driven Krylov extension rounds instead of one launch per iteration')
    ap.add_argument('--kmax', type=int, default=0, help='recycled interface basis capacity (vectors)')
    ap.add_argument('--mc', type=int, default=None)
    ap.add_argument('--block', type=int, default=None)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-budget', type=float, default=30.0)
    ap.add_argument("--prep-workers", type=int, default=8, help='serve: host preparation threads')
    ap.add_argument('--no-extras', action='store_true',
                    help='plain200k: skip the converged 2k-cell parity check, time-to-solution, batched and '
                         'served legs (kernel profiling runs)')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='run one 1 GiB device-to-device copy (known bytes) for HBM counter calibration')
    ap.add_argument('--no-fourier', action='store_true',
                    help='start from DeltaT = 0 (BTE_RTA with fourier=None); for kernel profiling')
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; multi-GPU runs (mode sharding + one '
                                   'all-reduce per outer iteration) are not part of it', 'n_gpus': args.gpus}))
        return 2

    if args.config in BATCH_CONFIGS:
        return run_batch(args)
    if args.config in SERVE_CONFIGS:
        return run_serve(args)
    if args.config in INVERSE_CONFIGS:
        return run_inverse(args)

    import torch
    from openbte_amd.bte import BTEDevice, default_config
    from openbte_amd.operator import build_host_operator
    from openbte_amd.fourier import Fourier

    geo, mat, bcs, eff = make_case(args.config)
    n, M, L = geo.n_elems, mat.grid[0], int(np.prod(mat.grid[1:]))
    print(f'# config {args.config}: n={n} M={M} L={L}', file=sys.stderr, flush=True)
    cpu = None
    c1 = None
    extras = args.config == 'plain200k' and not args.no_extras
    if args.config in CONFIGS3D and not args.no_cpu:
        # one SuperLU factorisation of a 3D system of this size takes minutes
        print('# no CPU leg for the 3D configuration (SuperLU fill-in)', file=sys.stderr, flush=True)
    elif not args.no_cpu:
        # CPU legs first, before this process touches the GPU (their workers fork)
        from oracle import fourier_oracle as fo
        if extras:
            c1 = c1_oracle(host_cores()[0])
            print(f'# oracle 2k-cell converged solve: {c1["n_iter"]} iterations, {c1["seconds"]:.1f}s, '
                  f'kappa {c1["kappa"]:.10f}', file=sys.stderr, flush=True)
        T_cpu = fo.fourier(geo, mat.thermal_conductivity, bcs, eff).variables['Temperature_Fourier']['data']
        cpu = cpu_baseline(geo, mat, bcs, eff, T_cpu, budget_s=args.cpu_budget)
        if c1 is not None:
            cpu['converged_2k_value'] = c1['updates_per_s']
            cpu['converged_2k_sample'] = (f'2k-cell workload ({c1["n"]} cells, 30 MFP x 48 directions) solved in '
                                          f'full by the oracle on {c1["workers"]} forked workers: '
                                          f'{c1["n_iter"]} outer iterations to a kappa change < 1e-6, '
                                          f'{c1["seconds"]:.2f}s wall incl. Fourier and factorisations')
        print(f'# cpu baseline {cpu["value"]:.3e} updates/s on {cpu["cores"]} cores', file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    fourier_info = None
    if args.no_fourier:
        T0 = np.zeros(n)
    else:
        torch.cuda.synchronize()
        tf0 = time.perf_counter()
        f = Fourier(geo, mat.thermal_conductivity, bcs, eff, verbose=False)
        T0 = f.variables['Temperature_Fourier']['data']
        tf = time.perf_counter() - tf0
        fourier_info = {'seconds': tf, 'iterations': f.aux['n_iter'],
                        'cg_iterations': [c[0] for c in f.aux['cg']],
                        'kappa_diffusive': f.aux['kappa_diffusive']}
        print(f'# fourier done ({f.aux["n_iter"]} iterations, {tf:.2f}s)', file=sys.stderr, flush=True)
    op = build_host_operator(geo, mat, bcs, eff)
    cfg = default_config([op])  # the shape BTE_RTA takes for this problem (mc 8 at 200k cells)
    cfg.window = not args.no_window
    if args.mc:
        cfg.mc = args.mc
    if args.block:
        cfg.block = args.block
    if args.kmax:
        cfg.kmax = args.kmax
    dev = BTEDevice(op, cfg=cfg)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    print(f'# setup {t_setup:.1f}s, cuts={dev.n_cut}', file=sys.stderr, flush=True)

    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device='cuda')  # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    # warmup: throwaway solve
    dev.reset(T0)
    for _ in range(args.warmup):
        dev.outer_iteration()
    torch.cuda.synchronize()

    # timed: outer iterations 1..K of a fresh solve
    dev.reset(T0)
    torch.cuda.synchronize()
    timer = []
    exts = []
    norms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        exts.append(dev.outer_iteration(timer=timer))
        norms.append(dev.norm2.clone())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nrm = [float(np.linalg.norm(T0))] + [float(np.sqrt(x.item())) for x in norms]
    errors = [abs(nrm[i + 1] - nrm[i]) / nrm[i + 1] for i in range(args.steps)]
    kappa = float(dev.buf['kappa'].sum().item())
    updates = n * M * L * args.steps
    value = updates / dt

    by_mode = {m: [a.elapsed_time(b) * 1e-3 for (mm, a, b) in timer if mm == m] for m in (0, 1, 2)}
    mean_ms = {f'mode{m}': (float(np.mean(v)) * 1e3 if v else None) for m, v in by_mode.items()}
    total_ms = {f'mode{m}': float(np.sum(v)) * 1e3 for m, v in by_mode.items()}
    # roofline of the final full sweep (mode 2): algorithmic bytes = 28 B per
    # cell*mode update for triangles, 32 B for tetrahedra (source read 8 +
    # solution write 8 + 8 per upwind gather, 1.5 / 2 inflow faces on average)
    # x the n*M*L updates of one launch.  Its compulsory HBM bytes are reported
    # beside it: records read once, the interface inflows, and either one
    # write of every value or, when the sweep forms the mode sums from its LDS
    # window, the per-chunk sums (dim + 1 doubles per cell and chunk; the
    # values later levels read from t come on top)
    per_update = 28.0 if geo.dim == 2 else 32.0
    sweep_bytes = float(n * M * L) * per_update
    fused = dev.window_positions() > 0 and 'part_chunk' in dev.buf and not dev.plan.keep_t
    if fused:
        compulsory = float(L * n * 64 + L * dev.n_chunks * n * (geo.dim + 1) * 8 + dev.n_cut * M * 8)
        kernel = 'k_sweep final sweep (mode 2: wall temperatures, kappa, per-chunk mode sums from the LDS window)'
    else:
        compulsory = float(L * n * 64 + L * n * M * 8 + dev.n_cut * M * 8)
        kernel = 'k_sweep final sweep (mode 2: every solution value written)'
    t_full = float(np.mean(by_mode[2]))
    achieved = sweep_bytes / t_full
    window_positions, n_cut, kmax = dev.window_positions(), dev.n_cut, dev.cfg.kmax
    print(f'# {args.steps} steps in {dt:.3f}s: {value:.3e} updates/s; extensions per step {exts}; '
          f'sweep ms mean by mode {mean_ms}, total {total_ms}; kappa {kappa:.6f}',
          file=sys.stderr, flush=True)

    traffic = None
    tf = os.path.join(ROOT, 'profiles', f'k_sweep_{args.config}_traffic.json')
    if os.path.exists(tf):
        rec = json.load(open(tf))
        if rec.get('mc') == cfg.mc and rec.get('block') == cfg.block:
            f_, w_ = rec['fetch_kib_per_dispatch'], rec['write_kib_per_dispatch']
            traffic = 1024.0 * (f_['mode2'] + w_['mode2'])

    north = None
    if extras:
        del dev
        north = north_star_legs(args, geo, mat, bcs, eff, c1, value)
    frac_counter = traffic / t_full / HBM_PEAK if traffic else None
    out = {
        'metric': BASELINE_METRIC,
        'metric_detail': ('value: cell*mode updates/s of the plain solve, outer iterations 1..K from the Fourier '
                          'guess, every system to direct-solver accuracy; latency, batched/served ratios and '
                          'the kappa error vs the oracle in north_star'),
        'value': value,
        'unit': 'cell*mode updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': dt / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': ('synthetic (seeded jittered-Delaunay pore cell, seeded lognormal Si-like material)'
                 if geo.dim == 2 else 'synthetic (seeded tetrahedral porous slab, seeded lognormal Si-like material)'),
        'config': {'workload': (f'{args.config}: L={CONFIGS[args.config][0]:g}nm porosity 0.25, n={n} cells, '
                                if args.config in CONFIGS else
                                f'{args.config}: periodic porous slab, n={n} tetrahedra, ')
                               + f'{M} MFP x {L} directions = {M * L} systems',
                   'mc': cfg.mc, 'block': cfg.block, 'lds_window_positions': window_positions,
                   'kappa_eff_after_steps': kappa,
                   'extensions_per_step': exts, 'n_cut_faces': n_cut,
                   'outer_error_per_step': errors, 'kmax': kmax, 'gcr_rtol': cfg.rtol,
                   'fourier': fourier_info, 'setup_s': t_setup},
        'roofline': {'bound': 'hbm', 'achieved': achieved / 1e9, 'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK, 'traffic': traffic,
                     'achieved_model': (f'algorithmic bytes: {per_update:g} B per cell*mode update (source 8 + '
                                        f'solution 8 + 8 per upwind gather) x n*M*L updates of one final-sweep '
                                        f'launch / its mean launch time (HIP events on the launch stream)'),
                     'traffic_source': ('stored rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE per dispatch of this '
                                        'kernel and shape (profiles/); not measured in this run')
                     if traffic else None,
                     'frac_counter': frac_counter,
                     'frac_end_to_end': value * per_update / HBM_PEAK,
                     'limiter': ('CU pipelines, not HBM: the windowed sweep keeps upwind values in LDS, so its '
                                 'HBM traffic is below the model bytes; counters show VALU ~0.7 and the '
                                 'texture data path ~0.9 busy'),
                     'kernel': kernel,
                     'bytes_per_launch': sweep_bytes, 'bytes_per_update': per_update,
                     'compulsory_bytes_per_launch': compulsory, 'mean_launch_s': t_full,
                     'sweep_ms_mean_by_mode': mean_ms, 'sweep_ms_total_by_mode': total_ms},
        'cpu_baseline': cpu,
    }
    if north is not None:
        if cpu is not None:
            north['gpu_over_cpu_steady'] = value / cpu['value']
            north['gpu_over_cpu_first_iteration'] = value / cpu['first_iteration_value']
        out['north_star'] = north
    print(json.dumps(out))
    return 0


if __name__ == '__main__':
    sys.exit(main())
