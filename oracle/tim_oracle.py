"""Interfacial-transmission (TIM) BTE model on the CPU (TEST INFRASTRUCTURE).

The oracle of the differentiation path; only tests, the smoke check and the
benchmark's checker leg may import it.

Model.  Pixel densities rho (cells inherit the density of their pixel) map to
a transmission per internal side between different pixels by the
reference's interpolation (openbte/inverse/bte.py:153-156):

    rho~ = k0 + rho (1 - k0),  k0 = 1e-12,   t = 2 rho~_a rho~_b / (rho~_a + rho~_b)

and t = 1 inside a pixel.  Across a side with transmission t the inflow of a
cell is t times the upwind value (the periodic jump included) plus (1 - t)
times the diffuse re-emission of the cell's own outflow through that side,
i.e. each side of such an interface is a diffuse wall of weight 1 - t whose
temperature is lagged like the walls of openbte/bte.py:73-78,199-214 (the
reference's inverse operator reflects the same way, openbte/inverse/bte.py:
160-176, with the reflection kept implicit).  The kappa contact flux through a
side is scaled by its t (openbte/inverse/bte.py:54-55).  With t = 1 everywhere
the model is exactly the plain solve of ``bte_oracle``.

Functions:
  side_transmission -- the interpolation above and its derivatives
  assemble_tim      -- ``bte_oracle.Operators`` with transmission applied and
                       the reflecting sides appended to the diffuse list
  iterate           -- Jacobi outer iteration (openbte/bte.py:193-256)
  dense_kappa       -- the converged fixed point as one dense linear solve
                       (tiny meshes): pins ``iterate`` and, by central
                       differences, the adjoint gradient of the GPU path

Parity pinning: t = 1 reduces to ``bte_oracle`` (itself pinned); ``iterate``
equals ``dense_kappa`` at convergence; the gradient is checked against
central differences of ``dense_kappa`` (exact linear algebra, no iteration
error).
"""
from __future__ import annotations

from typing import NamedTuple, Optional

import numpy as np
import scipy.sparse.linalg as spla

from oracle import bte_oracle as bo

K0 = 1e-12


def side_transmission(geo, pixel_of_cell: np.ndarray, rho: np.ndarray):
    """Transmission of every internal side (``geo.internal`` order) and its
    derivatives with respect to the densities of the two pixels.

    Returns (t, dt_da, dt_db, pa, pb, interface) where pa/pb are the pixels of
    side_elem_map[:,0]/[:,1] and ``interface`` marks sides between two
    different pixels (t = 1 and zero derivatives elsewhere)."""
    inn = np.asarray(geo.internal)
    sem = geo.side_elem_map[inn]
    pa = pixel_of_cell[sem[:, 0]]
    pb = pixel_of_cell[sem[:, 1]]
    interface = pa != pb
    r = K0 + np.asarray(rho, np.float64) * (1.0 - K0)
    a, b = r[pa], r[pb]
    t = np.where(interface, 2 * a * b / (a + b), 1.0)
    dt_da = np.where(interface, 2 * b * b / (a + b) ** 2 * (1.0 - K0), 0.0)
    dt_db = np.where(interface, 2 * a * a / (a + b) ** 2 * (1.0 - K0), 0.0)
    return t, dt_da, dt_db, pa, pb, interface


class TimOps(NamedTuple):
    ops: bo.Operators      # transmission applied; diffuse list = walls + reflecting sides
    weight: np.ndarray     # (nd,) injection weight: 1 for walls, 1 - t for reflecting sides
    n_wall: int            # number of real diffuse wall sides (first in the list)
    c_t: np.ndarray        # (n_contact,) transmission of the contact sides


def assemble_tim(geo, mat, bcs, eff, t_int: np.ndarray, interface: np.ndarray) -> TimOps:
    """``bte_oracle.assemble`` with transmissions ``t_int`` (one per internal
    side, ``geo.internal`` order); ``interface`` selects the sides that get
    reflecting walls."""
    ops = bo.assemble(geo, mat, bcs, eff)
    dim, n = ops.dim, ops.n_elems
    inn = np.asarray(geo.internal)
    nint = len(inn)
    t_dir = np.concatenate([t_int, t_int])          # directed faces (i->j, then j->i)
    Gm = ops.Gm * t_dir[None, :]
    # periodic jump: the transmitted part carries the jump
    s = np.asarray(mat.vmfp, np.float64)[:, :dim]
    sem = geo.side_elem_map
    i, j = sem[:, 0], sem[:, 1]
    side_t = np.ones(len(geo.sides))
    side_t[inn] = t_int
    P = np.zeros((ops.n_dirs, n))
    for region, jump in bcs.periodic.items():
        sides = np.asarray(geo.periodic_sides[region])
        tmp = s @ (geo.normal_areas[sides] / geo.elem_volumes[i[sides]][:, None]).T
        np.add.at(P.T, i[sides], -(np.clip(tmp, None, 0) * jump * side_t[sides]).T)
        tmp = s @ (-geo.normal_areas[sides] / geo.elem_volumes[j[sides]][:, None]).T
        np.add.at(P.T, j[sides], (np.clip(tmp, None, 0) * jump * side_t[sides]).T)
    # the plain operator's periodic part is replaced (mixed/nano parts kept)
    P_plain = np.zeros_like(P)
    for region, jump in bcs.periodic.items():
        sides = np.asarray(geo.periodic_sides[region])
        tmp = s @ (geo.normal_areas[sides] / geo.elem_volumes[i[sides]][:, None]).T
        np.add.at(P_plain.T, i[sides], -(np.clip(tmp, None, 0) * jump).T)
        tmp = s @ (-geo.normal_areas[sides] / geo.elem_volumes[j[sides]][:, None]).T
        np.add.at(P_plain.T, j[sides], (np.clip(tmp, None, 0) * jump).T)
    pert = ops.perturbation - P_plain + P
    # reflecting sides: both sides of every interface, db out of the cell
    sel = np.nonzero(interface)[0]
    cell = np.concatenate([i[inn[sel]], j[inn[sel]]])
    db = np.concatenate([geo.normal_areas[inn[sel]] / geo.elem_volumes[i[inn[sel]]][:, None],
                         -geo.normal_areas[inn[sel]] / geo.elem_volumes[j[inn[sel]]][:, None]])
    w = 1.0 - np.concatenate([t_int[sel], t_int[sel]])
    Gb = np.einsum('mqj,sj->mqs', mat.sigma[:, :, :dim], db)
    tot = np.clip(Gb, None, 0).sum(axis=(0, 1))
    R = np.divide(1.0, tot, out=np.zeros_like(tot), where=tot != 0)
    GG = np.concatenate([ops.GG, np.clip(Gb, 0, None) * R[None, None, :]], axis=2)
    Gbm2 = np.concatenate([ops.Gbm2, np.clip(s @ db.T, None, 0)], axis=1)
    i_diffuse = np.concatenate([ops.i_diffuse, cell])
    weight = np.concatenate([np.ones(len(ops.i_diffuse)), w])
    # contact flux scaled by the contact side's transmission
    c_t = np.ones(len(ops.e1))
    if ops.has_kappa and len(ops.e1):
        contact = np.asarray(geo.periodic_sides[eff.contact])
        c_t = side_t[contact]
    out = ops._replace(Gm=Gm, perturbation=pert, i_diffuse=i_diffuse, GG=GG, Gbm2=Gbm2)
    return TimOps(out, weight, len(ops.i_diffuse), c_t)


def _rhs(T: TimOps, l, m, DeltaT, TB):
    ops = T.ops
    lam = ops.mfps[m]
    B = DeltaT + lam * ops.perturbation[l] + ops.heat
    if len(ops.i_diffuse):
        np.add.at(B, ops.i_diffuse, lam * (-TB * ops.Gbm2[l] * T.weight))
    return B


def _kappa_ml(T: TimOps, l, m, x):
    ops = T.ops
    sn = ops.sigma_normal[m, l]
    k = np.dot(x[ops.e1] * T.c_t, np.clip(sn, 0, None))
    if len(ops.e2):
        k += np.dot(x[ops.e2] * T.c_t, np.clip(sn, None, 0))
    k += ops.alpha * np.sum(T.c_t * np.clip(sn, None, 0))
    return k * ops.normalization


def iterate(T: TimOps, DeltaT0: np.ndarray, max_iter: int = 200, rtol: float = 1e-12):
    """Jacobi outer iteration of the TIM model; returns (kappa, DeltaT, TB, n_iter)."""
    ops = T.ops
    DeltaT = DeltaT0.copy()
    TB = DeltaT[ops.i_diffuse].copy()
    lus = {}
    norm_old = np.linalg.norm(DeltaT)
    kappa = 0.0
    for it in range(1, max_iter + 1):
        dt = np.zeros(ops.n_elems)
        tb = np.zeros(len(ops.i_diffuse))
        kappa = 0.0
        for l in range(ops.n_dirs):
            for m in range(ops.n_mfp):
                if (l, m) not in lus:
                    lus[(l, m)] = spla.splu(bo.system_matrix(ops, l, m))
                x = lus[(l, m)].solve(_rhs(T, l, m, DeltaT, TB))
                dt += ops.t_coeff[m, l] * x
                tb += -x[ops.i_diffuse] * ops.GG[m, l]
                kappa += _kappa_ml(T, l, m, x)
        DeltaT, TB = dt, tb
        nrm = np.linalg.norm(DeltaT)
        err = abs(nrm - norm_old) / nrm
        norm_old = nrm
        if err <= rtol:
            break
    return kappa, DeltaT, TB, it


def dense_kappa(T: TimOps) -> float:
    """Converged kappa of the TIM model from one dense solve of the coupled
    system in (T_ml for all modes) (tiny meshes only)."""
    ops = T.ops
    n, M, L = ops.n_elems, ops.n_mfp, ops.n_dirs
    nd = len(ops.i_diffuse)
    N = n * M * L
    if N > 6000:
        raise ValueError('dense_kappa is for tiny problems')
    A = np.zeros((N, N))
    f = np.zeros(N)
    idx = lambda l, m: (l * M + m) * n  # noqa: E731
    for l in range(L):
        for m in range(M):
            o = idx(l, m)
            A[o:o + n, o:o + n] = bo.system_matrix(ops, l, m).toarray()
            f[o:o + n] = _rhs(T, l, m, np.zeros(n), np.zeros(nd))
    # coupling through DeltaT = sum t T and TB = -sum GG T[i_d]
    for l in range(L):
        for m in range(M):
            o = idx(l, m)
            lam = ops.mfps[m]
            for l2 in range(L):
                for m2 in range(M):
                    o2 = idx(l2, m2)
                    A[o:o + n, o2:o2 + n] -= ops.t_coeff[m2, l2] * np.eye(n)
                    if nd:
                        # row i_d of mode (l,m) gets lam * w * (-TB_d) Gbm2 with TB_d = -sum GG T[i_d]
                        coef = lam * T.weight * ops.Gbm2[l] * ops.GG[m2, l2]   # (nd,)
                        np.add.at(A, (o + ops.i_diffuse, o2 + ops.i_diffuse), -coef)
    x = np.linalg.solve(A, f)
    kappa = 0.0
    for l in range(L):
        for m in range(M):
            o = idx(l, m)
            kappa += _kappa_ml(T, l, m, x[o:o + n])
    return kappa


def reduced_kappa(T: TimOps) -> float:
    """Converged kappa of the TIM model from an exact solve of the outer
    iteration's fixed point reduced to its coupling variables x = (DeltaT,
    TB): every mode's solution is T_ml = A_ml^-1 (rhs_ml(0) + R_ml x), so
    x = M x + f with M and f formed from one multi-right-hand-side SuperLU
    solve per mode, and (I - M) x = f is solved densely (least squares: the
    periodic problem fixes DeltaT only up to a constant, which kappa does not
    see).  No iteration error; scales to full mode grids on small meshes
    (pinned against ``dense_kappa`` on tiny ones)."""
    ops = T.ops
    n, M, L = ops.n_elems, ops.n_mfp, ops.n_dirs
    nd = len(ops.i_diffuse)
    N = n + nd
    Mx = np.zeros((N, N))
    f = np.zeros(N)
    zero_n, zero_d = np.zeros(n), np.zeros(nd)
    lus = {}
    for l in range(L):
        for m in range(M):
            lu = lus[(l, m)] = spla.splu(bo.system_matrix(ops, l, m))
            lam = ops.mfps[m]
            R = np.zeros((n, N))
            R[:, :n] = np.eye(n)
            if nd:
                # d rhs / d TB_s = lam * w_s * (-Gbm2_ls) at cell i_d[s] (accumulated)
                np.add.at(R, (ops.i_diffuse, n + np.arange(nd)), lam * (-ops.Gbm2[l] * T.weight))
            X = lu.solve(R)
            x0 = lu.solve(_rhs(T, l, m, zero_n, zero_d))
            Mx[:n] += ops.t_coeff[m, l] * X
            f[:n] += ops.t_coeff[m, l] * x0
            if nd:
                Mx[n:] += -ops.GG[m, l][:, None] * X[ops.i_diffuse]
                f[n:] += -ops.GG[m, l] * x0[ops.i_diffuse]
    x = np.linalg.lstsq(np.eye(N) - Mx, f, rcond=None)[0]
    DeltaT, TB = x[:n], x[n:]
    kappa = 0.0
    for l in range(L):
        for m in range(M):
            kappa += _kappa_ml(T, l, m, lus[(l, m)].solve(_rhs(T, l, m, DeltaT, TB)))
    return kappa
