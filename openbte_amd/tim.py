"""Differentiable kappa_eff of pixel densities (interfacial-transmission model).

Model (the reference's inverse-design interpolation, openbte/inverse/bte.py:
150-176): pixel densities rho in [0, 1] map to a transmission per side
between two different pixels,

    rho~ = k0 + rho (1 - k0),  k0 = 1e-12,   t = 2 rho~_a rho~_b / (rho~_a + rho~_b),

and t = 1 inside a pixel.  Across such a side a cell receives t times the
upwind value (the periodic jump included) plus (1 - t) times the diffuse
re-emission of its own outflow through the side: each side of an interface is
a diffuse wall of weight 1 - t with a lagged wall temperature, exactly like the
walls of openbte/bte.py:73-78,199-214.  The kappa contact flux through a side
is scaled by its t.  With t = 1 everywhere the model is the plain solve.

Adjoint.  For upwind finite volumes (V A_s)^T = V A_{-s} (D_l is the same for
s and -s because every cell is closed), so A_ml^{-T} = V A_{m,lbar}^{-1} V^{-1}
with lbar the reversed direction.  Writing the adjoint as lambda_ml =
t_ml V mu'_ml, the transposed coupled system (mode sums DeltaT, TB) becomes the
forward outer iteration on mu' -- the same operator, the same sweeps in
reversed level order, the same GCR interface solve and reductions -- with
one constant source, the kappa functional's derivative, in place of the
perturbation.  The face derivatives are then one HIP reduction over modes
(obte_tim_face_grad); the chain rule to pixels is host work.

``KappaTIM`` wraps forward and adjoint in a torch.autograd.Function.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _lib
from .bte import BTEDevice, SweepConfig
from .operator import HostOperator, build_host_operator, build_schedule

K0 = 1e-12


def side_transmission(geo, pixel_of_cell: np.ndarray, rho: np.ndarray):
    """Transmission t of every internal side (``geo.internal`` order) and its
    derivatives with respect to the densities of the pixels of
    ``side_elem_map[:,0]`` (a) and ``[:,1]`` (b) (openbte/inverse/bte.py:153-157)."""
    inn = np.asarray(geo.internal)
    sem = np.asarray(geo.side_elem_map)[inn]
    pa = pixel_of_cell[sem[:, 0]]
    pb = pixel_of_cell[sem[:, 1]]
    interface = pa != pb
    r = K0 + np.asarray(rho, np.float64) * (1.0 - K0)
    a, b = r[pa], r[pb]
    t = np.where(interface, 2.0 * a * b / (a + b), 1.0)
    dt_da = np.where(interface, 2.0 * b * b / (a + b) ** 2 * (1.0 - K0), 0.0)
    dt_db = np.where(interface, 2.0 * a * a / (a + b) ** 2 * (1.0 - K0), 0.0)
    return t, dt_da, dt_db, pa, pb, interface


def reversed_directions(dirs: np.ndarray) -> np.ndarray:
    """lbar[l] with s_lbar = -s_l; raises if the direction set is not closed
    under reversal."""
    d = np.asarray(dirs, np.float64)
    dot = d @ d.T
    lbar = np.argmin(dot, axis=1)
    if np.abs(d[lbar] + d).max() > 1e-9:
        raise ValueError('the direction set has no reversed partner for every direction')
    return lbar.astype(np.int32)


def _with_interfaces(op: HostOperator, geo, interface: np.ndarray) -> tuple:
    """Host operator whose wall list also holds one reflecting side per
    directed interface face (weights set later), and the face tables."""
    inn = np.asarray(geo.internal)
    nint = len(inn)
    F = 2 * nint
    face_itf = np.concatenate([interface, interface])
    faces = np.nonzero(face_itf)[0]                    # directed faces with a reflecting side
    n_real = len(op.i_diffuse)
    side_of_face = -np.ones(F, np.int32)
    side_of_face[faces] = n_real + np.arange(len(faces))
    cell = op.i_mat[faces].astype(np.int64)
    db = op.k_face[faces]
    Gb = np.einsum('mlj,sj->mls', op.sigma, db)
    tot = np.clip(Gb, None, 0).sum(axis=(0, 1))
    rinv = np.divide(1.0, tot, out=np.zeros_like(tot), where=tot != 0)
    # merge the reflecting entries into the per-cell wall lists
    n = op.n
    old_cell = np.repeat(np.arange(n), np.diff(op.wall_ptr))
    all_cell = np.concatenate([old_cell, cell])
    all_side = np.concatenate([op.wall_side.astype(np.int64), -2 - (n_real + np.arange(len(faces)))])
    all_db = np.concatenate([op.wall_db.reshape(-1, op.dim), db])
    all_temp = np.concatenate([op.wall_temp, np.zeros(len(faces))])
    o = np.argsort(all_cell, kind='stable')
    wall_ptr = np.zeros(n + 1, np.int64)
    np.add.at(wall_ptr, all_cell + 1, 1)
    op2 = op._replace(
        wall_ptr=np.cumsum(wall_ptr).astype(np.int32), wall_side=all_side[o].astype(np.int32),
        wall_db=np.ascontiguousarray(all_db[o]), wall_temp=all_temp[o],
        i_diffuse=np.concatenate([op.i_diffuse, cell]),
        wall_dbs=np.ascontiguousarray(np.concatenate([op.wall_dbs.reshape(-1, op.dim), db])),
        wall_rinv=np.concatenate([op.wall_rinv, rinv]),
        t_face=np.ones(F), c_t=np.ones(len(op.ce1)))
    # position of each reflecting side's entry in the merged wall list
    inv = np.empty(len(o), np.int64)
    inv[o] = np.arange(len(o))
    entry_of_side = inv[len(old_cell):]
    return op2, faces, side_of_face, entry_of_side


class TIMSolver:
    """Forward kappa_eff(rho) and its adjoint gradient on one GPU.

    ``pixel_of_cell`` maps cells to pixels (n_pixels = max + 1).  Forward and
    adjoint solves find the fixed point of the outer iteration of
    openbte/bte.py:193-256 to a relative residual ``rtol`` (far tighter than
    the reference's stop rule, so that gradients and finite differences are
    meaningful), with the interface Krylov solve at ``cfg.rtol`` (default here
    1e-13).  Successive solves start from the previous solution.

    The gradient needs every forward mode solution T_ml next to the adjoint
    ones.  By default (``store_forward=True``) the forward solve keeps a copy
    of its solution workspace (L x M x n doubles).  With
    ``store_forward=False`` it keeps only its converged state (DeltaT, TB)
    and interface values, and the gradient recomputes the forward and adjoint
    solutions ``dir_block`` directions at a time by one final sweep each:
    extra memory ``dir_block / L`` of a workspace instead of a whole one, for
    two extra sweeps in total."""

    def __init__(self, geo, mat, bcs, eff, pixel_of_cell: np.ndarray, cfg: Optional[SweepConfig] = None,
                 rtol: float = 1e-12, max_iter: int = 3000, restart: int = 80, store_forward: bool = True,
                 dir_block: int = 8):
        if not torch.cuda.is_available():
            raise RuntimeError('openbte_amd.TIMSolver needs a ROCm GPU; there is no CPU fallback')
        self.geo = geo
        self.pixel_of_cell = np.asarray(pixel_of_cell, np.int64)
        self.n_pixels = int(self.pixel_of_cell.max()) + 1
        self.rtol, self.max_iter = float(rtol), int(max_iter)
        op = build_host_operator(geo, mat, bcs, eff)
        self.c_sides = np.asarray(geo.periodic_sides[eff.contact])
        if op.contact_kind != 1:
            raise ValueError('the differentiable kappa needs a periodic kappa contact')
        _, _, _, _, _, interface = side_transmission(geo, self.pixel_of_cell, np.ones(self.n_pixels))
        self.interface = interface
        self.jump0 = op.jump_face.copy()
        op, faces, side_of_face, entry = _with_interfaces(op, geo, interface)
        self.op, self.faces, self.entry = op, faces, entry
        self.lbar = reversed_directions(op.dirs)
        # adjoint requirements: t and sigma symmetric under reversal, sigma = C t Lambda s
        if np.abs(op.t_coeff[:, self.lbar] - op.t_coeff).max() > 1e-12 * np.abs(op.t_coeff).max():
            raise ValueError('t_coeff is not symmetric under direction reversal')
        tl = op.t_coeff[:, :, None] * op.mfps[:, None, None] * op.dirs[None, :, :]
        mask = np.abs(tl) > 1e-14 * np.abs(tl).max()
        ratio = op.sigma[mask] / tl[mask]
        self.c_sigma = float(np.median(ratio))
        if np.abs(ratio / self.c_sigma - 1).max() > 1e-9:
            raise ValueError('sigma_ml is not proportional to t_ml Lambda_m s_l')
        self.sched = build_schedule(op)
        cfg = cfg or SweepConfig()
        if cfg.rtol > 1e-13:
            cfg = SweepConfig(**{**vars(cfg), 'rtol': 1e-13})
        self.dev = BTEDevice(op, sched=self.sched, cfg=cfg)
        d = self.dev
        d.keep_solutions(True)  # the gradient reads the forward and adjoint solutions from t
        dv = d.dev
        self.vol = torch.as_tensor(np.asarray(geo.elem_volumes, np.float64), device=dv)
        self.i_face = torch.as_tensor(op.i_mat, dtype=torch.int32, device=dv)
        self.j_face = torch.as_tensor(op.j_mat, dtype=torch.int32, device=dv)
        self.side_of_face = torch.as_tensor(side_of_face, dtype=torch.int32, device=dv)
        self.jump0_d = torch.as_tensor(self.jump0, dtype=torch.float64, device=dv)
        self.lbar_d = torch.as_tensor(self.lbar, dtype=torch.int32, device=dv)
        self.heat0 = torch.zeros_like(d.buf['heat'])
        self.nd = len(op.i_diffuse)
        self.restart = int(restart)
        self.store_forward = bool(store_forward)
        self.dir_block = max(1, int(dir_block))
        self.x_fwd = self.x_adj = None
        self.t_fwd = self.y_fwd = None
        self.n_iter = 0
        self.kappa = None
        self.rho = None

    # -- transmissions ---------------------------------------------------------
    def set_density(self, rho: np.ndarray) -> None:
        rho = np.asarray(rho, np.float64).reshape(self.n_pixels)
        t, da, db, pa, pb, _ = side_transmission(self.geo, self.pixel_of_cell, rho)
        self.rho, self.t_int, self.dt_da, self.dt_db, self.pa, self.pb = rho, t, da, db, pa, pb
        d, op = self.dev, self.op
        t_face = np.concatenate([t, t])
        d.buf['t_face'].copy_(torch.as_tensor(t_face))
        d.buf['jump_face'].copy_(torch.as_tensor(self.jump0 * t_face))
        wt = op.wall_temp.copy()
        wt[self.entry] = 1.0 - t_face[self.faces]
        d.buf['wall_temp'].copy_(torch.as_tensor(wt))
        side_t = np.ones(len(self.geo.sides))
        side_t[np.asarray(self.geo.internal)] = t
        d.buf['c_t'].copy_(torch.as_tensor(side_t[self.c_sides]))
        _lib.call('obte_bte_assemble', d.plan, d.stream)

    # -- solves ----------------------------------------------------------------
    # The converged state x = (DeltaT, TB) is the fixed point of the outer
    # iteration x -> G(x) = M x + f (openbte/bte.py:193-256; one evaluation =
    # sources, sweeps, interface solve, mode sums).  Near-opaque interfaces make
    # the spectral radius of M approach 1, so the fixed point is found by
    # restarted GMRES on (I - M) x = f with G as the operator; the limit is the
    # same as the Jacobi iteration's.
    def _set_state(self, x: torch.Tensor) -> None:
        d, n = self.dev, self.op.n
        d.buf['dt'][d.cur].copy_(x[:n])
        if self.nd:
            d.buf['tb'][d.cur][:self.nd].copy_(x[n:])

    def _G(self, x: torch.Tensor) -> torch.Tensor:
        d = self.dev
        self._set_state(x)
        d.outer_iteration()
        self.evals += 1
        return torch.cat([d.buf['dt'][d.cur], d.buf['tb'][d.cur][:self.nd]])

    def _solve(self, x0: Optional[torch.Tensor]) -> torch.Tensor:
        d = self.dev
        n, N = self.op.n, self.op.n + self.nd
        d.reset(np.zeros(n))
        self.evals = 0
        zero = torch.zeros(N, dtype=torch.float64, device=d.dev)
        f = self._G(zero)
        fn = float(torch.linalg.norm(f).item())
        x = zero.clone() if x0 is None else x0.clone()
        if fn == 0.0:
            return zero
        m = self.restart
        V = torch.zeros(m + 1, N, dtype=torch.float64, device=d.dev)
        while self.evals < self.max_iter:
            r = f - (x - (self._G(x) - f)) if x.abs().max() > 0 else f.clone()
            beta = float(torch.linalg.norm(r).item())
            if beta <= self.rtol * fn:
                break
            V[0] = r / beta
            H = np.zeros((m + 1, m))
            cs, sn = np.zeros(m), np.zeros(m)
            g = np.zeros(m + 1)
            g[0] = beta
            k = 0
            for j in range(m):
                w = V[j] - (self._G(V[j]) - f)
                # classical Gram-Schmidt, twice
                h = V[:j + 1] @ w
                w = w - V[:j + 1].T @ h
                h2 = V[:j + 1] @ w
                w = w - V[:j + 1].T @ h2
                hj = (h + h2).cpu().numpy()
                hn = float(torch.linalg.norm(w).item())
                H[:j + 1, j] = hj
                H[j + 1, j] = hn
                if hn > 0:
                    V[j + 1] = w / hn
                for i in range(j):  # previous Givens rotations
                    t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                    H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                    H[i, j] = t
                den = np.hypot(H[j, j], H[j + 1, j])
                cs[j], sn[j] = (1.0, 0.0) if den == 0 else (H[j, j] / den, H[j + 1, j] / den)
                H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
                H[j + 1, j] = 0.0
                g[j + 1] = -sn[j] * g[j]
                g[j] = cs[j] * g[j]
                k = j + 1
                if abs(g[j + 1]) <= self.rtol * fn or hn == 0 or self.evals >= self.max_iter:
                    break
            y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
            x = x + V[:k].T @ torch.as_tensor(y, device=d.dev)
        # leave the device state, workspace and mode sums at the solution
        self._G(x)
        return x

    def forward(self, rho: np.ndarray) -> float:
        self.set_density(rho)
        d = self.dev
        self.x_fwd = self._solve(self.x_fwd)
        self.n_iter = self.evals
        self.kappa = float(d.buf['kappa'].sum().item())
        if self.store_forward:
            self.t_fwd = d.buf['t'].clone()
        else:
            # the final sweep's interface values: with the converged state they
            # reproduce every forward solution exactly (obte_bte_sweep mode 2)
            self.t_fwd = None
            self.y_fwd = d.gbuf['y'].clone()
        self.tb_fwd = self.x_fwd[self.op.n:].clone()
        return self.kappa

    def _adjoint_source(self) -> torch.Tensor:
        """Constant adjoint source per (direction d, position): the kappa
        functional's derivative for the reversed direction, divided by t_ml
        Lambda_m V (openbte/bte.py:237-243 with the contact transmission)."""
        op, geo = self.op, self.geo
        n, L = op.n, op.n_dirs
        na = np.asarray(geo.normal_areas)[self.c_sides]
        side_t = np.ones(len(geo.sides))
        side_t[np.asarray(geo.internal)] = self.t_int
        ct = side_t[self.c_sides]
        g = op.dirs @ na.T                                   # (L, nc) s_d . nA
        src = np.zeros((L, n))
        scale = op.normalization * self.c_sigma
        # reversed direction: (s_l . nA)^+ = -(s_d . nA)^-, (s_l . nA)^- = -(s_d . nA)^+
        np.add.at(src.T, op.ce1, (-np.clip(g, None, 0) * ct * scale).T)
        np.add.at(src.T, op.ce2, (-np.clip(g, 0, None) * ct * scale).T)
        src /= np.asarray(geo.elem_volumes)[None, :]
        pos_src = np.take_along_axis(src, self.sched.order.reshape(L, n).astype(np.int64), axis=1)
        return torch.as_tensor(pos_src.reshape(-1), device=self.dev.dev)

    def adjoint(self) -> None:
        d = self.dev
        P = d.plan
        self.adj_src = self._adjoint_source()
        d.buf['pert'].copy_(self.adj_src)
        heat_ptr = P.heat
        P.heat = self.heat0.data_ptr()
        try:
            self.x_adj = self._solve(self.x_adj)
            self.n_iter_adj = self.evals
            self.y_adj = d.gbuf['y'].clone()
        finally:
            P.heat = heat_ptr
            _lib.call('obte_bte_assemble', P, d.stream)     # restores the forward perturbation

    def _resweep(self, x: torch.Tensor, y: torch.Tensor, dirs: np.ndarray, adjoint: bool) -> None:
        """Recompute the final-sweep solutions of directions ``dirs`` (all MFP
        chunks) into the workspace from a converged state ``x`` and interface
        values ``y``; the other directions' rows are left as they are."""
        d = self.dev
        P = d.plan
        active = torch.zeros(self.op.n_dirs, d.n_chunks, dtype=torch.int32)
        active[torch.as_tensor(np.asarray(dirs, np.int64))] = 1
        active = active.to(d.dev)
        heat_ptr = P.heat
        if adjoint:
            d.buf['pert'].copy_(self.adj_src)
            P.heat = self.heat0.data_ptr()
        try:
            self._set_state(x)
            _lib.call('obte_bte_source', P, d.stream)
            _lib.call('obte_bte_sweep', P, 2, y.data_ptr(), d.gbuf['x'].data_ptr(), active.data_ptr(), 0, d.stream)
        finally:
            if adjoint:
                P.heat = heat_ptr
                _lib.call('obte_bte_assemble', P, d.stream)

    def gradient(self) -> np.ndarray:
        """d kappa_eff / d rho for the densities of the last ``forward``."""
        if self.kappa is None:
            raise RuntimeError('call forward first')
        self.adjoint()
        d, op = self.dev, self.op
        F = len(op.i_mat)
        grad_f = torch.zeros(F, dtype=torch.float64, device=d.dev)
        args = (self.lbar_d.data_ptr(), self.i_face.data_ptr(), self.j_face.data_ptr(), self.side_of_face.data_ptr(),
                self.tb_fwd.data_ptr(), self.jump0_d.data_ptr(), self.vol.data_ptr(), F, grad_f.data_ptr(), d.stream)
        if self.store_forward:
            _lib.call('obte_tim_face_grad', d.plan, self.t_fwd.data_ptr(), d.buf['t'].data_ptr(), *args)
            contact = self._contact_term(self.t_fwd, 0, op.n_dirs)
        else:
            # forward and adjoint solutions recomputed a block of directions at
            # a time: forward rows of the block are copied out, then the
            # adjoint rows of the reversed block are swept in place
            L = op.n_dirs
            rows = (op.n + d.ncut_max) * d.n_chunks * d.cfg.mc
            tw = d.buf['t'].view(L, rows)
            blk = torch.empty(min(self.dir_block, L), rows, dtype=torch.float64, device=d.dev)
            contact = np.zeros(F // 2)
            for l0 in range(0, L, self.dir_block):
                l1 = min(L, l0 + self.dir_block)
                self._resweep(self.x_fwd, self.y_fwd, np.arange(l0, l1), adjoint=False)
                blk[:l1 - l0].copy_(tw[l0:l1])
                self._resweep(self.x_adj, self.y_adj, self.lbar[l0:l1], adjoint=True)
                _lib.call('obte_tim_face_grad_dirs', d.plan, blk.data_ptr(), l0, l1, int(l0 > 0),
                          d.buf['t'].data_ptr(), *args)
                contact += self._contact_term(blk[:l1 - l0], l0, l1)
        gf = grad_f.cpu().numpy()
        nint = F // 2
        g_side = gf[:nint] + gf[nint:]
        g_side += contact
        grad = np.zeros(self.n_pixels)
        np.add.at(grad, self.pa, g_side * self.dt_da)
        np.add.at(grad, self.pb, g_side * self.dt_db)
        return grad

    def _contact_term(self, t_fwd: torch.Tensor, l0: int, l1: int) -> np.ndarray:
        """d kappa / d t of the contact sides through the flux formula itself,
        from the forward solutions of directions l0 <= l < l1 (``t_fwd`` holds
        those directions only)."""
        op, geo, d = self.op, self.geo, self.dev
        n, M, mc, nch = op.n, op.n_mfp, d.cfg.mc, d.n_chunks
        L = l1 - l0
        rows = n + d.ncut_max
        T = t_fwd.reshape(L, nch, rows, mc)
        pos = torch.as_tensor(self.sched.pos.reshape(op.n_dirs, n)[l0:l1].astype(np.int64), device=d.dev)
        e1 = torch.as_tensor(op.ce1.astype(np.int64), device=d.dev)
        e2 = torch.as_tensor(op.ce2.astype(np.int64), device=d.dev)
        li = torch.arange(L, device=d.dev)[:, None]
        T1 = T[li, :, pos[:, e1], :]                    # (L, nc, nch, mc)
        T2 = T[li, :, pos[:, e2], :]
        T1 = T1.reshape(L, len(op.ce1), nch * mc)[:, :, :M].cpu().numpy()
        T2 = T2.reshape(L, len(op.ce1), nch * mc)[:, :, :M].cpu().numpy()
        na = np.asarray(geo.normal_areas)[self.c_sides]
        sn = np.einsum('mlj,cj->lcm', op.sigma[:, l0:l1], na)      # (L, nc, M)
        per = (np.clip(sn, 0, None) * T1 + np.clip(sn, None, 0) * (T2 + op.contact_jump)).sum(axis=(0, 2))
        out = np.zeros(len(op.i_mat) // 2)
        side_index = {s: k for k, s in enumerate(np.asarray(geo.internal))}
        for c, s in enumerate(self.c_sides):
            out[side_index[int(s)]] += op.normalization * per[c]
        return out


class KappaTIM(torch.autograd.Function):
    """kappa = KappaTIM.apply(rho, solver): forward solve on the GPU, backward
    by the adjoint solve (``TIMSolver.gradient``)."""

    @staticmethod
    def forward(ctx, rho: torch.Tensor, solver: TIMSolver):
        k = solver.forward(rho.detach().cpu().numpy().astype(np.float64))
        ctx.solver = solver
        return torch.tensor(k, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.solver.gradient()
        return grad_out * torch.as_tensor(g, dtype=torch.float64, device=grad_out.device), None
