"""CPU restatement of the reference's steady-state BTE solve (TEST INFRASTRUCTURE).

This module is the oracle: FP64 NumPy + SciPy SuperLU, written from the
mathematics of openbte/bte.py:19-271.  Only the tests, the smoke check and the
CPU-baseline leg of the benchmark may import it; the product path never does.

Pinning (how each function is checked, see tests/test_oracle_*.py):
  assemble          -- dense brute force on tiny meshes (one with cells that
                       have two diffuse walls), face-closure and
                       zero-net-flux invariants, closed form on uniform quads
  rhs               -- dense fixed point on the two-wall mesh (the wall
                       source accumulates over a cell's walls); isothermal
                       source against a side-by-side assembly, uniform
                       temperature between two contacts at the same T0;
                       the heat term heat_source * heat_source_coeff added
                       to every system (openbte/bte.py:199-214) by energy
                       balance: the generated heat V H 1e-18 leaves through
                       the isothermal contacts, and every column of sides
                       passes on the heat generated behind it
                       (tests/test_oracle_energy.py)
  solve_system      -- dense numpy.linalg.solve of the same matrix
  bte_rta           -- closed form kappa on uniform quads (exact upwind
                       solution), dense fixed point of the coupled system,
                       mirror symmetry, independence of the worker split

Semantics kept from the reference:
  * per-(l, m) operator A = I + Lambda_m (diag(D_l) + G^-_l) on the shared
    pattern, duplicates summed (openbte/bte.py:181-185)
  * Jacobi outer iteration: every system of iteration k uses DeltaT and TB of
    iteration k-1; the flux J uses the old DeltaT (openbte/bte.py:196-243)
  * stop rule |‖DeltaT‖ - ‖DeltaT_old‖| / ‖DeltaT‖ <= bte_rtol or
    bte_max_iter iterations (openbte/bte.py:193-196,251-254)
"""
from __future__ import annotations

import multiprocessing as mp
import os
import time
from typing import NamedTuple, Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class Operators(NamedTuple):
    n_elems: int
    n_mfp: int
    n_dirs: int
    dim: int
    mfps: np.ndarray          # (M,)
    t_coeff: np.ndarray       # (M, L)
    sigma: np.ndarray         # (M, L, dim_sigma)
    i_mat: np.ndarray         # (K,) row cell of each directed face
    j_mat: np.ndarray         # (K,) neighbour cell
    Gm: np.ndarray            # (L, K)   min(s_l . k_f, 0)
    D: np.ndarray             # (L, n)   sum of outflow coefficients
    i_diffuse: np.ndarray     # (nd,)
    GG: np.ndarray            # (M, L, nd)
    Gbm2: np.ndarray          # (L, nd)
    perturbation: np.ndarray  # (L, n)
    heat: np.ndarray          # (n,)  heat_source * heat_source_coeff
    im: np.ndarray
    jm: np.ndarray
    # effective conductivity
    has_kappa: bool
    normalization: float
    e1: np.ndarray
    e2: np.ndarray
    alpha: float
    sigma_normal: np.ndarray  # (M, L, n_contact)
    side_areas: np.ndarray    # (n_contact,)
    is_nanowire: bool


def _n_dirs(mat) -> int:
    g = mat.grid
    return int(g[1] * g[2]) if len(g) == 3 else int(g[1])


def assemble(geo, mat, bcs, effective_thermal_conductivity=None, heat_source=None) -> Operators:
    """Upwind finite-volume operators (openbte/bte.py:47-158)."""
    dim = geo.dim
    n = geo.n_elems
    M = int(mat.grid[0])
    L = _n_dirs(mat)
    s = np.asarray(mat.vmfp, dtype=np.float64)
    sem = geo.side_elem_map
    i = sem[:, 0]
    j = sem[:, 1]
    inn = np.asarray(geo.internal)

    # directed internal faces (openbte/bte.py:48-54)
    k_mat = np.concatenate([geo.normal_areas[inn] / geo.elem_volumes[i[inn]][:, None],
                            -geo.normal_areas[inn] / geo.elem_volumes[j[inn]][:, None]])
    i_mat = np.concatenate([i[inn], j[inn]])
    j_mat = np.concatenate([j[inn], i[inn]])

    # G = s . k, split in inflow/outflow parts; D = sum of outflow (openbte/bte.py:58-62)
    G = s[:, :dim] @ k_mat.T
    Gp = np.clip(G, 0, None)
    Gm = np.clip(G, None, 0)
    D = np.zeros((L, n))
    np.add.at(D.T, i_mat, Gp.T)

    # diffuse walls (openbte/bte.py:68-87)
    if len(bcs.diffuse) > 0:
        s_diffuse = np.asarray(geo.side_physical_regions[bcs.diffuse])
        i_diffuse = i[s_diffuse]
        db = geo.normal_areas[s_diffuse] / geo.elem_volumes[i_diffuse][:, None]
        Gb = np.einsum('mqj,sj->mqs', mat.sigma[:, :, :dim], db)
        Gbp = np.clip(Gb, 0, None)
        Gbm = np.clip(Gb, None, 0)
        tot = Gbm.sum(axis=(0, 1))
        R = np.divide(1.0, tot, out=np.zeros_like(tot), where=tot != 0)
        GG = Gbp * R[None, None, :]
        sd = s[:, :dim] @ db.T
        Gbp2 = np.clip(sd, 0, None)
        Gbm2 = np.clip(sd, None, 0)
        np.add.at(D.T, i_diffuse, Gbp2.T)
    else:
        i_diffuse = np.zeros(0, np.int64)
        GG = np.zeros((M, L, 0))
        Gbm2 = np.zeros((L, 0))

    # isothermal (mixed) walls; boundary_conductance is not used by the BTE (openbte/bte.py:90-96)
    RHS_ISO = np.zeros((L, n))
    for key, contact in bcs.mixed.items():
        sides = np.asarray(geo.side_physical_regions[key])
        tmp = geo.normal_areas[sides] / geo.elem_volumes[i[sides]][:, None]
        tmp = s @ tmp.T if s.shape[1] == dim else s[:, :dim] @ tmp.T
        np.add.at(RHS_ISO.T, i[sides], -(np.clip(tmp, None, 0) * contact['value']).T)
        np.add.at(D.T, i[sides], np.clip(tmp, 0, None).T)

    # periodic jumps (openbte/bte.py:99-110)
    P = np.zeros((L, n))
    for region, jump in bcs.periodic.items():
        sides = np.asarray(geo.periodic_sides[region])
        tmp = geo.normal_areas[sides] / geo.elem_volumes[i[sides]][:, None]
        tmp = s[:, :dim] @ tmp.T
        np.add.at(P.T, i[sides], -(np.clip(tmp, None, 0) * jump).T)
        tmp = -geo.normal_areas[sides] / geo.elem_volumes[j[sides]][:, None]
        tmp = s[:, :dim] @ tmp.T
        np.add.at(P.T, j[sides], (np.clip(tmp, None, 0) * jump).T)

    P_nano = -s[:, 2] if bcs.is_nanowire else np.zeros(L)
    perturbation = P + RHS_ISO + P_nano[:, None]

    im = np.concatenate([i_mat, np.arange(n)])
    jm = np.concatenate([j_mat, np.arange(n)])

    if heat_source is None:
        hs = np.zeros(n)
    elif hasattr(heat_source, 'variables'):
        hs = np.asarray(heat_source.variables['Heat_Source']['data'], dtype=np.float64)
    else:
        hs = np.asarray(heat_source, dtype=np.float64)
    heat = hs * mat.heat_source_coeff

    # effective-conductivity contact (openbte/bte.py:132-158)
    has_kappa = effective_thermal_conductivity is not None
    e1 = e2 = np.zeros(0, np.int64)
    alpha = 0.0
    sigma_normal = np.zeros((M, L, 0))
    side_areas = np.zeros(0)
    normalization = 0.0
    if has_kappa:
        contact = effective_thermal_conductivity.contact
        normalization = float(effective_thermal_conductivity.normalization)
        if bcs.is_nanowire:
            e1 = np.arange(n)
            sigma_normal = np.einsum('ml,c->mlc', mat.sigma[:, :, 2], geo.elem_volumes)
        else:
            bc = bcs.get_bc_by_name(contact)
            if bc == 'Mixed':
                sides = np.asarray(geo.side_physical_regions[contact])
                e1 = sem[sides, 0]
                alpha = float(bcs.mixed[contact]['value'])
            else:
                sides = np.asarray(geo.periodic_sides[contact])
                e1 = sem[sides, 0]
                e2 = sem[sides, 1]
                alpha = float(bcs.periodic[contact])
            sigma_normal = np.einsum('mli,ci->mlc', mat.sigma[:, :, :dim], geo.normal_areas[sides])
            side_areas = np.linalg.norm(geo.normal_areas[sides], axis=1)

    return Operators(n, M, L, dim, np.asarray(mat.mfps, np.float64), np.asarray(mat.t_coeff, np.float64),
                     np.asarray(mat.sigma, np.float64), i_mat, j_mat, Gm, D, i_diffuse, GG, Gbm2,
                     perturbation, heat, im, jm, has_kappa, normalization, e1, e2, alpha,
                     sigma_normal, side_areas, bool(bcs.is_nanowire))


def system_matrix(ops: Operators, l: int, m: int) -> sp.csc_matrix:
    """A_ml = I + Lambda_m (diag(D_l) + G^-_l), duplicates summed (openbte/bte.py:181-185)."""
    lam = ops.mfps[m]
    data = np.concatenate([lam * ops.Gm[l], lam * ops.D[l] + 1.0])
    return sp.csc_matrix((data, (ops.im, ops.jm)), shape=(ops.n_elems, ops.n_elems), dtype=np.float64)


def rhs(ops: Operators, l: int, m: int, DeltaT: np.ndarray, TB: np.ndarray) -> np.ndarray:
    """B_ml = DeltaT + Lambda_m pert_l + heat; diffuse walls add
    Lambda_m (-TB Gbm2_l) with accumulation on multi-wall cells (openbte/bte.py:199-214)."""
    lam = ops.mfps[m]
    B = DeltaT + lam * ops.perturbation[l] + ops.heat
    if len(ops.i_diffuse):
        np.add.at(B, ops.i_diffuse, lam * (-TB * ops.Gbm2[l]))
    return B


class ModeSums(NamedTuple):
    DeltaT: np.ndarray
    TB: np.ndarray
    J: np.ndarray
    K: np.ndarray
    S: np.ndarray


def accumulate(ops: Operators, l: int, m: int, T: np.ndarray, DeltaT_old: np.ndarray, acc: ModeSums) -> None:
    """Mode-sum contributions of one solved system (openbte/bte.py:222-243)."""
    acc.DeltaT[:] += ops.t_coeff[m, l] * T
    if len(ops.i_diffuse):
        acc.TB[:] += -T[ops.i_diffuse] * ops.GG[m, l]
    acc.J[:] += np.outer(T - DeltaT_old, ops.sigma[m, l])
    if ops.has_kappa:
        sn = ops.sigma_normal[m, l]
        if ops.is_nanowire:
            acc.K[m, l] = np.dot(T[ops.e1], sn) * ops.normalization
        else:
            acc.K[m, l] = np.dot(T[ops.e1], np.clip(sn, 0, None)) * ops.normalization
            acc.S[m, l] = np.dot(T[ops.e1], ops.side_areas) * ops.normalization
        if len(ops.e2) > 0:
            acc.K[m, l] += np.dot(T[ops.e2], np.clip(sn, None, 0)) * ops.normalization
        acc.K[m, l] += ops.alpha * np.sum(np.clip(sn, None, 0)) * ops.normalization


def _empty_sums(ops: Operators) -> ModeSums:
    return ModeSums(np.zeros(ops.n_elems), np.zeros(len(ops.i_diffuse)),
                    np.zeros((ops.n_elems, ops.sigma.shape[2])),
                    np.zeros((ops.n_mfp, ops.n_dirs)), np.zeros((ops.n_mfp, ops.n_dirs)))


class _Worker:
    """Direction slice of the outer iteration with a per-worker LU cache
    (the reference keeps up to 1e3 factorisations per process, openbte/bte.py:177)."""

    def __init__(self, ops: Operators, dirs: np.ndarray, cache_lu: bool = True):
        self.ops = ops
        self.dirs = dirs
        self.cache_lu = cache_lu
        self.lu = {}

    def factor(self, l, m):
        key = (l, m)
        if key in self.lu:
            return self.lu[key]
        f = spla.splu(system_matrix(self.ops, l, m))
        if self.cache_lu:
            self.lu[key] = f
        return f

    def iterate(self, DeltaT: np.ndarray, TB: np.ndarray) -> ModeSums:
        ops = self.ops
        acc = _empty_sums(ops)
        for l in self.dirs:
            for m in range(ops.n_mfp):
                B = rhs(ops, l, m, DeltaT, TB)
                T = self.factor(l, m).solve(B)
                accumulate(ops, l, m, T, DeltaT, acc)
        return acc


def _reduce(parts):
    out = parts[0]
    for p in parts[1:]:
        for a, b in zip(out, p):
            a += b
    return out


def _worker_main(conn, ops, dirs):
    w = _Worker(ops, dirs)
    while True:
        msg = conn.recv()
        if msg is None:
            break
        DeltaT, TB = msg
        conn.send(tuple(w.iterate(DeltaT, TB)))
    conn.close()


class OuterLoop:
    """Outer (source) iteration of openbte/bte.py:193-256 over ``n_workers``
    forked processes, each owning ``np.array_split(directions, n_workers)``
    (openbte/utils.py:209-211).  Partial sums are reduced in worker order."""

    def __init__(self, ops: Operators, n_workers: int = 1):
        self.ops = ops
        self.n_workers = max(1, int(n_workers))
        self.slices = [s for s in np.array_split(np.arange(ops.n_dirs), self.n_workers)]
        self.procs = []
        self.conns = []
        if self.n_workers > 1:
            ctx = mp.get_context('fork')
            for sl in self.slices:
                a, b = ctx.Pipe()
                p = ctx.Process(target=_worker_main, args=(b, ops, sl), daemon=True)
                p.start()
                self.procs.append(p)
                self.conns.append(a)
        else:
            self.local = _Worker(ops, self.slices[0])

    def step(self, DeltaT, TB) -> ModeSums:
        if self.n_workers == 1:
            return self.local.iterate(DeltaT, TB)
        for c in self.conns:
            c.send((DeltaT, TB))
        parts = [ModeSums(*c.recv()) for c in self.conns]
        return _reduce(parts)

    def close(self):
        for c in self.conns:
            c.send(None)
        for p in self.procs:
            p.join()
        self.procs, self.conns = [], []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def bte_rta(geo, mat, bcs, fourier=None, effective_thermal_conductivity=None, heat_source=None,
            direct: bool = True, bte_max_iter: int = 100, bte_rtol: float = 1e-4,
            n_workers: int = 1, verbose: bool = False, history: Optional[list] = None):
    """Restatement of ``BTE_RTA`` (openbte/bte.py:19-271) returning the same
    ``SolverResults`` layout.  ``direct=False`` in the reference builds its
    lgmres matrix from the direction index (openbte/bte.py:188); that path is
    not restated and ``direct`` is ignored."""
    from openbte_amd.objects import SolverResults
    ops = assemble(geo, mat, bcs, effective_thermal_conductivity, heat_source)
    if fourier is None:
        DeltaT = np.zeros(geo.n_elems)
    else:
        DeltaT = np.array(fourier.variables['Temperature_Fourier']['data'], dtype=np.float64)
    TB = DeltaT[ops.i_diffuse].copy() if len(ops.i_diffuse) else np.zeros(0)
    J = np.zeros((geo.n_elems, ops.sigma.shape[2]))
    K = np.zeros((ops.n_mfp, ops.n_dirs))
    S = np.zeros_like(K)
    norm_old = np.linalg.norm(DeltaT)
    error, n_iter = 1.0, 0
    with OuterLoop(ops, n_workers) as loop:
        while error > bte_rtol and n_iter < bte_max_iter:
            n_iter += 1
            sums = loop.step(DeltaT, TB)
            DeltaT, TB, J, K, S = sums
            norm_T = np.linalg.norm(DeltaT)
            error = abs(norm_T - norm_old) / norm_T
            norm_old = norm_T
            if history is not None:
                history.append((n_iter, float(K.sum()), float(error)))
            if verbose:
                print("Kappa (BTE): {:E}, Iterations {:n}, Error {:E}".format(K.sum(), n_iter, error))
    kappa_eff = float(np.sum(K))
    variables = {'Temperature_BTE': {'data': DeltaT, 'units': 'K'},
                 'Flux_BTE': {'data': J * 1e9, 'units': 'W/m/m'}}
    return SolverResults(kappa_eff=kappa_eff, variables=variables,
                         aux={'suppression': S, 'kappa_eff': kappa_eff, 'n_iter': n_iter})


def time_outer_iterations(geo, mat, bcs, fourier, eff, n_workers: int, dirs: Optional[np.ndarray] = None,
                          n_iter: int = 2):
    """CPU-baseline timer: wall clock of outer iterations restricted to the
    direction subset ``dirs`` (all MFP bins each), split over ``n_workers``
    forked processes.  Returns a list of per-iteration seconds; iteration 1
    includes the LU factorisations, later ones reuse the cached factors."""
    ops = assemble(geo, mat, bcs, eff)
    if dirs is not None:
        ops = ops._replace(n_dirs=ops.n_dirs)
    DeltaT = np.array(fourier.variables['Temperature_Fourier']['data'], dtype=np.float64) if fourier else np.zeros(geo.n_elems)
    TB = DeltaT[ops.i_diffuse].copy()
    sel = np.arange(ops.n_dirs) if dirs is None else np.asarray(dirs)
    slices = np.array_split(sel, max(1, n_workers))
    times = []
    ctx = mp.get_context('fork')
    conns, procs = [], []
    for sl in slices:
        a, b = ctx.Pipe()
        p = ctx.Process(target=_worker_main, args=(b, ops, sl), daemon=True)
        p.start()
        conns.append(a)
        procs.append(p)
    try:
        for _ in range(n_iter):
            t0 = time.perf_counter()
            for c in conns:
                c.send((DeltaT, TB))
            parts = [ModeSums(*c.recv()) for c in conns]
            times.append(time.perf_counter() - t0)
    finally:
        for c in conns:
            c.send(None)
        for p in procs:
            p.join()
    return times
