"""``BTE_RTA`` on MI355X: the steady-state mode-resolved BTE solve
(drop-in for openbte/bte.py:19-271).

Per outer iteration (the reference's Jacobi source iteration,
openbte/bte.py:193-256) every (direction l, MFP bin m) system
A_ml T = B_ml with A_ml = I + Lambda_m (D_l + G^-_l) is solved exactly by

  1. a level-scheduled upwind sweep with the periodic inflow faces cut
     (source B_ml, interface values y from the previous iteration),
  2. GCR on the cut-face interface equations y = R S(B, y), reusing the
     basis kept from earlier iterations (the transfer operator does not
     change between iterations), extending it only while residuals exceed
     the tolerance,
  3. a final sweep with the converged y that also accumulates the mode sums
     (lattice temperature, wall temperatures, flux, kappa and suppression).

All of this runs as HIP kernels from libopenbte_amd; this module only owns
the device buffers (torch tensors) and drives the loop.
"""
from __future__ import annotations

import time
import warnings
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .objects import SolverResults
from .operator import (HostOperator, Schedule, build_host_operator, build_schedule, fixed_width_inflow,
                       level_pair_max, max_inflow, separable_velocities)


@dataclass
class SweepConfig:
    """Launch and Krylov knobs (picked per workload by ``openbte_amd.tune``)."""
    mc: int = 8              # MFP bins per workgroup chunk (looped by each lane)
    block: int = 256         # threads per sweep workgroup (one cell per lane per level)
    kmax: int = 128          # recycled GCR basis capacity per system (capped by basis_bytes)
    rtol: float = 1e-10      # interface residual tolerance
    max_extend: int = 64     # cap on basis extensions per outer iteration
    basis_bytes: float = 24e9  # HBM budget of the recycled basis (u and c)
    window: bool = True      # keep recent levels of a sweep in LDS where they fit
    wave_tasks: bool = True  # batches: one wave per task for problems with narrow levels
    fused_extend: bool = True  # one launch runs every task's whole Krylov extension loop
    separable_flux: bool = True  # one flux mode sum per position when the velocities allow it


class InterfaceNotConverged(UserWarning):
    """An outer iteration reached ``SweepConfig.max_extend`` interface Krylov
    extensions with systems still above the interface tolerance: its systems
    are not solved to direct-solver accuracy."""


def warn_unconverged(n_tasks: int, max_extend: int) -> None:
    warnings.warn(f'{n_tasks} (direction, chunk) tasks still above the interface tolerance after '
                  f'max_extend={max_extend} Krylov extensions in this outer iteration; raise '
                  f'SweepConfig.max_extend', InterfaceNotConverged, stacklevel=3)


@dataclass
class IterationStats:
    n_iter: int = 0
    sweeps: list = field(default_factory=list)       # system-sweeps per iteration (x M*L)
    extensions: list = field(default_factory=list)
    kappa: list = field(default_factory=list)
    error: list = field(default_factory=list)
    seconds: list = field(default_factory=list)


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError('openbte_amd.BTE_RTA needs a ROCm GPU; there is no CPU fallback')
    return torch.device('cuda', torch.cuda.current_device())


def _t(a, dtype, dev):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


def _kmax(cfg: SweepConfig, ny: int) -> int:
    """Krylov basis length: cfg.kmax capped so that the two basis blocks
    (u and c, kmax x ny doubles each) fit in cfg.basis_bytes."""
    return int(max(8, min(cfg.kmax, cfg.basis_bytes // (16 * ny))))


def part_stride(n: int, nf: int) -> int:
    """Doubles per direction of the per-direction sums (and per chunk of the
    per-chunk sums): n (nf + 1) rounded up to even, so rows are 16-byte
    aligned; nf = flux sums per position (``flux_sums``)."""
    return (n * (nf + 1) + 1) // 2 * 2


def flux_sums(op: HostOperator, cfg: SweepConfig):
    """``(nf, sig_a, sig_dir)``: flux mode sums per position and the separable
    factors of the group velocities (1 with them, else dim and None)."""
    a, d = separable_velocities(op.sigma) if cfg.separable_flux else (None, None)
    return (op.dim, None, None) if a is None else (1, a, d)


def _window_cap(cfg: SweepConfig, pair_max: int) -> int:
    """Positions of the LDS window of sweeps with this shape (0: none)."""
    return int(_lib.load().obte_sweep_window(int(bool(cfg.window)), int(pair_max), cfg.mc))


def device_bytes(op: HostOperator, sched: Schedule, cfg: Optional[SweepConfig] = None) -> int:
    """Device memory (bytes) a ``BTEDevice`` of this operator allocates; the
    serving loop admits a request only when this fits its HBM budget."""
    cfg = cfg or SweepConfig()
    n, L, M, dim = op.n, op.n_dirs, op.n_mfp, op.dim
    mc = cfg.mc
    nch = (M + mc - 1) // mc
    n_cut = int(sched.cut_off[-1])
    ncut_max = int(np.diff(sched.cut_off).max()) if L > 0 else 0
    kin = max(2, max_inflow(sched))
    n_wall = len(op.i_diffuse)
    one = lambda k: max(1, k)  # noqa: E731  (empty host arrays are stored as one element)
    i4 = (2 * L * n + len(sched.lvl_off) + len(sched.lvl_ptr) + 2 * L * kin * n + len(sched.cut_off)
          + len(sched.cut_exit) + len(op.cf_ptr) + len(op.cf_face) + len(op.wall_ptr) + one(len(op.wall_side))
          + one(n_wall) + one(len(op.ce1)) + one(len(op.ce2)))
    f8 = (op.dirs.size + op.k_face.size + op.jump_face.size + max(dim, op.wall_db.size) + one(len(op.wall_temp))
          + one(n_wall) + (op.wall_dbs.size if n_wall else dim) + max(dim, op.c_na.size) + op.mfps.size
          + op.t_coeff.size + op.sigma.size + op.heat.size)
    if op.t_face is not None:
        f8 += len(op.t_face) + max(1, len(op.c_t))          # t_face, c_t
    f8 += L * n + L * n * 8                                  # pert, rec
    f8 += 2 * n + 2 * one(n_wall) + n * dim                  # dt, tb, j
    f8 += L * nch * (n + ncut_max) * mc                      # t
    nf, sig_a, _ = flux_sums(op, cfg)
    if sig_a is not None:
        f8 += sig_a.size + L * dim                           # sig_a, sig_dir
    ps = part_stride(n, nf)
    f8 += L * ps + one(L * nch * n_wall)                     # part_tj, part_tb
    f8 += 2 * M * L + 4096 + 8                               # kappa, supp, scratch
    if _window_cap(cfg, level_pair_max(sched)) > 0:
        f8 += L * nch * ps                                   # part_chunk
    ny = max(1, nch * n_cut * mc)
    f8 += 4 * ny + 2 * _kmax(cfg, ny) * ny + M * L           # y r x w, u c, rnorm
    i4 += 2 * L * nch + L * nch * mc                         # k, active, lane_active
    return int(8 * f8 + 4 * i4)


class BTEDevice:
    """Device-resident operator, schedule and workspaces of one geometry and
    material.  Reusable across solves (warm state is reset by ``reset``)."""

    def __init__(self, op: HostOperator, sched: Optional[Schedule] = None, cfg: Optional[SweepConfig] = None,
                 device=None, n_active=None, norm2=None):
        self.op = op
        # a private copy: the basis capacity is clamped per problem below
        self.cfg = SweepConfig(**vars(cfg)) if cfg is not None else SweepConfig()
        self.dev = device or _dev()
        self.sched = sched if sched is not None else build_schedule(op)
        dev, f64, i32 = self.dev, torch.float64, torch.int32
        op, sc, cfg = self.op, self.sched, self.cfg
        n, L, M, dim = op.n, op.n_dirs, op.n_mfp, op.dim
        mc = cfg.mc
        nch = (M + mc - 1) // mc
        self.n_chunks = nch
        self.n_cut = int(sc.cut_off[-1])
        self.ncut_max = int(np.diff(sc.cut_off).max()) if L > 0 else 0
        self.lvl_pair_max = level_pair_max(sc)
        n_wall = len(op.i_diffuse)
        self.kin = max(2, max_inflow(sc))
        if self.kin > 3:
            raise ValueError(f'{self.kin} inflow faces per cell: only up to 3 are supported')
        nb_src, nb_face = fixed_width_inflow(sc, n, self.kin)
        # the sweep kernel indexes T rows with these values: check them here
        ncut_l = np.diff(sc.cut_off)
        if nb_src.size and (nb_src.min() < 0 or np.any(nb_src.max(axis=(1, 2)) >= n + ncut_l)):
            raise ValueError('inflow source row out of range')
        if len(sc.cut_exit) and self.n_cut and (sc.cut_exit[:self.n_cut].min() < 0
                                                 or sc.cut_exit[:self.n_cut].max() >= n):
            raise ValueError('cut exit position out of range')
        self.buf = b = {}
        # schedule
        b['order'] = _t(sc.order, i32, dev)
        b['pos'] = _t(sc.pos, i32, dev)
        b['lvl_off'] = _t(sc.lvl_off, i32, dev)
        b['lvl_ptr'] = _t(sc.lvl_ptr, i32, dev)
        b['nb_src'] = _t(nb_src, i32, dev)
        b['nb_face'] = _t(nb_face, i32, dev)
        b['cut_off'] = _t(sc.cut_off, i32, dev)
        b['cut_exit'] = _t(sc.cut_exit, i32, dev)
        # geometry / material
        b['dirs'] = _t(op.dirs, f64, dev)
        b['k_face'] = _t(op.k_face, f64, dev)
        b['jump_face'] = _t(op.jump_face, f64, dev)
        b['cf_ptr'] = _t(op.cf_ptr, i32, dev)
        b['cf_face'] = _t(op.cf_face, i32, dev)
        b['wall_ptr'] = _t(op.wall_ptr, i32, dev)
        ws = np.asarray(op.wall_side, np.int64)
        # diffuse index, -1 (isothermal) or -2 - k (reflecting interface side k): the
        # source kernel indexes TB with these, so check them here
        k = np.where(ws >= 0, ws, -2 - ws)
        if len(ws) and (np.any(k[ws != -1] >= max(n_wall, 0)) or np.any(k[ws != -1] < 0)):
            raise ValueError('wall entry refers to a diffuse side out of range')
        b['wall_side'] = _t(ws if len(ws) else np.zeros(1), i32, dev)
        b['wall_db'] = _t(op.wall_db if len(op.wall_db) else np.zeros((1, dim)), f64, dev)
        b['wall_temp'] = _t(op.wall_temp if len(op.wall_temp) else np.zeros(1), f64, dev)
        b['wall_cell'] = _t(op.i_diffuse if n_wall else np.zeros(1), i32, dev)
        b['wall_rinv'] = _t(op.wall_rinv if n_wall else np.zeros(1), f64, dev)
        b['wall_dbs'] = _t(op.wall_dbs if n_wall else np.zeros((1, dim)), f64, dev)
        b['ce1'] = _t(op.ce1 if len(op.ce1) else np.zeros(1), i32, dev)
        b['ce2'] = _t(op.ce2 if len(op.ce2) else np.zeros(1), i32, dev)
        b['c_na'] = _t(op.c_na if len(op.c_na) else np.zeros((1, dim)), f64, dev)
        b['mfps'] = _t(op.mfps, f64, dev)
        b['t_coeff'] = _t(op.t_coeff, f64, dev)
        b['sigma'] = _t(op.sigma, f64, dev)
        self.n_flux, sig_a, sig_dir = flux_sums(op, cfg)
        if sig_a is not None:
            b['sig_a'] = _t(sig_a, f64, dev)
            b['sig_dir'] = _t(sig_dir, f64, dev)
        b['heat'] = _t(op.heat, f64, dev)
        if op.t_face is not None:
            b['t_face'] = _t(op.t_face, f64, dev)
            b['c_t'] = _t(op.c_t if len(op.c_t) else np.zeros(1), f64, dev)
        # assembled operator
        b['pert'] = torch.zeros(L * n, dtype=f64, device=dev)
        b['rec'] = torch.zeros(L * n * 8, dtype=f64, device=dev)
        # state
        b['dt'] = [torch.zeros(n, dtype=f64, device=dev), torch.zeros(n, dtype=f64, device=dev)]
        b['tb'] = [torch.zeros(max(1, n_wall), dtype=f64, device=dev) for _ in range(2)]
        b['j'] = torch.zeros(n * dim, dtype=f64, device=dev)
        # zero-initialised: unused inflow slots read a finite value times 0
        b['t'] = torch.zeros(L * nch * (n + self.ncut_max) * mc, dtype=f64, device=dev)
        ps = part_stride(n, self.n_flux)
        b['part_tj'] = torch.empty(L * ps, dtype=f64, device=dev)
        b['part_tb'] = torch.empty(max(1, L * nch * n_wall), dtype=f64, device=dev)
        b['kappa'] = torch.zeros(M * L, dtype=f64, device=dev)
        b['supp'] = torch.zeros(M * L, dtype=f64, device=dev)
        b['scratch'] = torch.zeros(4096 + 8, dtype=f64, device=dev)
        if _window_cap(cfg, self.lvl_pair_max) > 0:
            # per-chunk mode sums of the windowed final sweep
            b['part_chunk'] = torch.empty(L * nch * ps, dtype=f64, device=dev)
        # interface Krylov state
        ny = max(1, nch * self.n_cut * mc)
        cfg.kmax = _kmax(cfg, ny)
        g = self.gbuf = {}
        g['y'] = torch.zeros(ny, dtype=f64, device=dev)
        g['r'] = torch.zeros(ny, dtype=f64, device=dev)
        g['x'] = torch.zeros(ny, dtype=f64, device=dev)
        g['w'] = torch.zeros(ny, dtype=f64, device=dev)
        g['u'] = torch.zeros(cfg.kmax * ny, dtype=f64, device=dev)
        g['c'] = torch.zeros(cfg.kmax * ny, dtype=f64, device=dev)
        g['k'] = torch.zeros(L * nch, dtype=i32, device=dev)
        g['active'] = torch.zeros(L * nch, dtype=i32, device=dev)
        g['lane_active'] = torch.zeros(L * nch * mc, dtype=i32, device=dev)
        # counters may live in a batch-wide array (one readback per iteration)
        g['n_active'] = n_active if n_active is not None else torch.zeros(1, dtype=i32, device=dev)
        g['rnorm'] = torch.zeros(M * L, dtype=f64, device=dev)
        self.ext_max = torch.zeros(1, dtype=i32, device=dev)  # extensions of the fused loop
        self.norm2 = norm2 if norm2 is not None else torch.zeros(1, dtype=f64, device=dev)
        self.cur = 0
        self.plan = self._make_plan()
        self.gcr = self._make_gcr()
        self.stream = torch.cuda.current_stream(self.dev).cuda_stream
        _lib.call('obte_bte_assemble', self.plan, self.stream)

    def nbytes(self) -> int:
        """Bytes of the device tensors this object owns (shared counters excluded)."""
        ts = [t for v in self.buf.values() for t in (v if isinstance(v, list) else [v])]
        ts += [t for k, t in self.gbuf.items() if k != 'n_active']
        return int(sum(t.numel() * t.element_size() for t in ts))

    def window_positions(self) -> int:
        """Positions of the LDS window the sweeps of this device use (0: none)."""
        return _window_cap(self.cfg, self.lvl_pair_max)

    def keep_solutions(self, on: bool = True) -> None:
        """Let final sweeps store every solution value in ``buf['t']`` (for
        callers that read it); by default a windowed final sweep stores only
        the values later positions read and forms the mode sums on chip."""
        self.plan.keep_t = int(bool(on))

    # -- C structs -------------------------------------------------------
    def _make_plan(self) -> _lib.BtePlan:
        op, b, cfg = self.op, self.buf, self.cfg
        P = _lib.BtePlan()
        P.n, P.n_dirs, P.n_mfp, P.dim = op.n, op.n_dirs, op.n_mfp, op.dim
        P.mc, P.n_chunks, P.block = cfg.mc, self.n_chunks, cfg.block
        P.n_cut, P.n_wall, P.n_contact, P.n_faces = self.n_cut, len(op.i_diffuse), len(op.ce1), len(op.i_mat)
        P.contact_kind, P.normalization, P.contact_jump = op.contact_kind, op.normalization, op.contact_jump
        P.kin, P.ncut_max = self.kin, self.ncut_max
        P.lvl_pair_max, P.sweep_window = self.lvl_pair_max, int(bool(cfg.window))
        for name in ('order', 'pos', 'lvl_off', 'lvl_ptr', 'nb_src', 'nb_face', 'cut_off',
                     'cut_exit', 'dirs', 'k_face', 'jump_face', 'cf_ptr', 'cf_face', 'wall_ptr', 'wall_side',
                     'wall_db', 'wall_temp', 'wall_cell', 'wall_rinv', 'wall_dbs', 'ce1', 'ce2', 'c_na', 'mfps',
                     't_coeff', 'sigma', 'heat', 'pert', 'rec', 't', 'part_tj',
                     'part_tb', 'kappa', 'supp', 'scratch'):
            setattr(P, name, b[name].data_ptr())
        P.nano = op.nano
        if 't_face' in b:
            P.t_face, P.c_t = b['t_face'].data_ptr(), b['c_t'].data_ptr()
        P.delta_t = b['dt'][self.cur].data_ptr()
        P.tb = b['tb'][self.cur].data_ptr()
        P.part_chunk = b['part_chunk'].data_ptr() if 'part_chunk' in b else 0
        P.keep_t = 0
        if 'sig_a' in b:
            P.sig_a, P.sig_dir = b['sig_a'].data_ptr(), b['sig_dir'].data_ptr()
        return P

    def _make_gcr(self) -> _lib.Gcr:
        g = self.gbuf
        G = _lib.Gcr()
        G.kmax, G.rtol = self.cfg.kmax, self.cfg.rtol
        for name in ('y', 'r', 'x', 'u', 'c', 'k', 'active', 'lane_active', 'w', 'n_active', 'rnorm'):
            setattr(G, name, g[name].data_ptr())
        return G

    # -- state -------------------------------------------------------------
    def reset(self, delta_t0: np.ndarray) -> None:
        """Initial lattice temperature; TB = DeltaT at the wall cells
        (openbte/bte.py:163-166); interface values and basis cleared."""
        self.cur = 0
        d0 = torch.as_tensor(np.asarray(delta_t0, dtype=np.float64)).to(self.dev)
        self.buf['dt'][0].copy_(d0)
        if len(self.op.i_diffuse):
            self.buf['tb'][0][:len(self.op.i_diffuse)].copy_(d0[torch.as_tensor(self.op.i_diffuse).to(self.dev)])
        self.plan.delta_t = self.buf['dt'][0].data_ptr()
        self.plan.tb = self.buf['tb'][0].data_ptr()
        _lib.call('obte_gcr_reset', self.plan, self.gcr, 1, self.stream)

    def _n_active(self) -> int:
        return int(self.gbuf['n_active'].item())

    def _sweep(self, mode, y, active, timer, lane_active=0):
        if timer is not None:
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
        _lib.call('obte_bte_sweep', self.plan, mode, y, self.gbuf['x'].data_ptr(), active, lane_active, self.stream)
        if timer is not None:
            ev1.record()
            timer.append((mode, ev0, ev1))

    def outer_iteration(self, max_extend: Optional[int] = None, timer: Optional[list] = None):
        """One Jacobi outer iteration; returns the number of basis extensions.
        ``timer`` (optional list) receives (mode, start, end) HIP events of every
        sweep launch, recorded on the launch stream."""
        P, G, st = self.plan, self.gcr, self.stream
        g = self.gbuf
        max_extend = self.cfg.max_extend if max_extend is None else max_extend
        _lib.call('obte_bte_source', P, st)
        ext = 0
        if self.n_cut > 0 and self.cfg.fused_extend:
            self._sweep(0, g['y'].data_ptr(), 0, timer)
            _lib.call('obte_gcr_project', P, G, st)
            if timer is not None:
                ev0 = torch.cuda.Event(enable_timing=True)
                ev1 = torch.cuda.Event(enable_timing=True)
                ev0.record()
            _lib.call('obte_gcr_extend', P, G, max_extend, self.ext_max.data_ptr(), st)
            if timer is not None:
                ev1.record()
                timer.append((1, ev0, ev1))
            ext = -1  # read once the rest of the iteration is queued
        elif self.n_cut > 0:
            self._sweep(0, g['y'].data_ptr(), 0, timer)
            _lib.call('obte_gcr_project', P, G, st)
            while ext < max_extend and self._n_active() > 0:
                _lib.call('obte_gcr_begin_extend', P, G, st)
                self._sweep(1, g['w'].data_ptr(), g['active'].data_ptr(), timer, g['lane_active'].data_ptr())
                _lib.call('obte_gcr_end_extend', P, G, st)
                ext += 1
            if ext >= max_extend and self._n_active() > 0:
                warn_unconverged(self._n_active(), max_extend)
        self._sweep(2, g['y'].data_ptr(), 0, timer)
        nxt = 1 - self.cur
        b = self.buf
        _lib.call('obte_bte_reduce', P, b['dt'][nxt].data_ptr(), b['j'].data_ptr(), b['tb'][nxt].data_ptr(),
                  self.norm2.data_ptr(), st)
        self.cur = nxt
        P.delta_t = b['dt'][nxt].data_ptr()
        P.tb = b['tb'][nxt].data_ptr()
        if ext < 0:
            ext = int(self.ext_max.item())
            if ext >= max_extend and self._n_active() > 0:
                warn_unconverged(self._n_active(), max_extend)
        return ext

    def solve(self, delta_t0: np.ndarray, bte_max_iter: int = 100, bte_rtol: float = 1e-4,
              verbose: bool = False, stats: Optional[IterationStats] = None, sync_each: bool = True):
        self.reset(delta_t0)
        norm_old = np.float64(np.linalg.norm(delta_t0))
        error, n_iter = 1.0, 0
        while error > bte_rtol and n_iter < bte_max_iter:
            n_iter += 1
            t0 = time.perf_counter()
            ext = self.outer_iteration()
            norm_t = np.sqrt(np.float64(self.norm2.item()))
            # NumPy division as in openbte/bte.py:251-252: a zero field gives
            # nan (not an exception), which ends the loop after this iteration
            with np.errstate(divide='ignore', invalid='ignore'):
                error = float(np.abs(norm_t - norm_old) / norm_t)
            norm_old = norm_t
            if stats is not None or verbose:
                kappa = float(self.buf['kappa'].sum().item())
                if stats is not None:
                    stats.n_iter = n_iter
                    stats.sweeps.append(2 + ext if self.n_cut > 0 else 1)
                    stats.extensions.append(ext)
                    stats.kappa.append(kappa)
                    stats.error.append(error)
                    stats.seconds.append(time.perf_counter() - t0)
                if verbose:
                    print("Kappa (BTE): {:E}, Iterations {:n}, Error {:E}".format(kappa, n_iter, error))
        return n_iter

    def results(self):
        op, b = self.op, self.buf
        K = b['kappa'].view(op.n_mfp, op.n_dirs).cpu().numpy().copy()
        S = b['supp'].view(op.n_mfp, op.n_dirs).cpu().numpy().copy()
        DeltaT = b['dt'][self.cur].cpu().numpy().copy()
        J = b['j'].view(op.n, op.dim).cpu().numpy().copy()
        return DeltaT, J, K, S


def default_config(ops: Sequence[HostOperator]) -> SweepConfig:
    """Launch shape of ``BTE_RTA`` / ``BTE_RTA_batch`` when none is given:
    chunks of 8 MFP bins, or of 4 when chunks of 8 leave fewer than two tasks
    (direction x chunk) per compute unit for some problem (measured: 30 bins
    x 48 directions +19%, 40 x 64 +13% at mc 4; 60 x 96 is faster at mc 8).
    The batched and single forms take the same rule, so their results stay
    equal bit for bit."""
    cfg = SweepConfig()
    cus = 256
    if torch.cuda.is_available():  # (without a GPU, BTEDevice raises)
        cus = getattr(torch.cuda.get_device_properties(torch.cuda.current_device()), 'multi_processor_count', cus)
    if min(op.n_dirs * ((op.n_mfp + 7) // 8) for op in ops) < 2 * cus:
        cfg.mc = 4
    return cfg


def BTE_RTA(geo, mat, bcs, fourier=None, effective_thermal_conductivity=None, heat_source=None,
            direct: bool = True, bte_max_iter: int = 100, bte_rtol: float = 1e-4,
            config: Optional[SweepConfig] = None, verbose: bool = True,
            stats: Optional[IterationStats] = None) -> SolverResults:
    """Drop-in for ``openbte.bte.BTE_RTA`` (openbte/bte.py:19-27): same
    arguments, same ``SolverResults`` (kappa_eff, aux{'suppression',
    'kappa_eff'}, variables{'Temperature_BTE', 'Flux_BTE'}).  ``direct`` is
    accepted for signature compatibility; every system is solved to
    direct-solver accuracy.  Like the reference (openbte/bte.py:256) it
    prints kappa and the outer error every iteration; ``verbose=False``
    silences that."""
    t0 = time.perf_counter()
    op = build_host_operator(geo, mat, bcs, effective_thermal_conductivity, heat_source)
    t1 = time.perf_counter()
    sched = build_schedule(op)
    t2 = time.perf_counter()
    if fourier is None:
        dt0 = np.zeros(geo.n_elems)
    else:
        dt0 = np.asarray(fourier.variables['Temperature_Fourier']['data'], dtype=np.float64)
    solver = BTEDevice(op, sched=sched, cfg=config or default_config([op]))
    torch.cuda.synchronize(solver.dev)
    t3 = time.perf_counter()
    n_iter = solver.solve(dt0, bte_max_iter=bte_max_iter, bte_rtol=bte_rtol, verbose=verbose, stats=stats)
    DeltaT, J, K, S = solver.results()
    t4 = time.perf_counter()
    timing = {'operator_s': t1 - t0, 'schedule_s': t2 - t1, 'device_setup_s': t3 - t2, 'iterations_s': t4 - t3}
    if op.contact_kind == 0:
        K = np.zeros_like(K)
        S = np.zeros_like(S)
    kappa_eff = float(np.sum(K))
    variables = {'Temperature_BTE': {'data': DeltaT, 'units': 'K'},
                 'Flux_BTE': {'data': J * 1e9, 'units': 'W/m/m'}}
    return SolverResults(kappa_eff=kappa_eff, variables=variables,
                         aux={'suppression': S, 'kappa_eff': kappa_eff, 'n_iter': n_iter, 'timing': timing})
