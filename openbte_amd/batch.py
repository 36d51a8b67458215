"""Many BTE solves on the device at once: one persistent sweep launch over
every (problem, direction, MFP chunk) task.

Each problem keeps its own operator, schedule, Krylov basis and state (a
``BTEDevice``); the batch adds device arrays of their plan structs and a task
table.  Tasks of one (problem, direction) share one of 8 XCD-local queues,
assigned greedily by cost (levels x cells) so that the queues balance, and
each queue is ordered by decreasing cost.  Every problem follows its own
outer iteration and stop rule (openbte/bte.py:193-256): a converged problem is
marked not live and every kernel skips it, so its result is the one a solve
of that problem alone produces.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .bte import BTEDevice, SweepConfig, default_config, warn_unconverged
from .objects import SolverResults
from .operator import batch_max_width, build_host_operator, build_schedule

N_QUEUES = 8
_CAPTURE_LOCK = threading.Lock()


def _struct_tensor(structs, dev) -> torch.Tensor:
    """Copy an array of ctypes structs to a device byte tensor."""
    raw = b''.join(bytes(s) for s in structs)
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(dev)


class BTEBatch:
    """A batch of problems sharing the launch shape (mc, block, kin)."""

    def __init__(self, ops: Sequence = (), cfg: Optional[SweepConfig] = None, scheds: Optional[Sequence] = None,
                 groups: Optional[int] = None, device=None, devices: Optional[Sequence[BTEDevice]] = None,
                 n_active: Optional[torch.Tensor] = None, norm2: Optional[torch.Tensor] = None,
                 slots: Optional[Sequence[int]] = None):
        """Build the batch from host operators (schedules of at most
        ``batch_max_width(mc, block)`` cells per level unless ``scheds`` are
        given), or
        (``devices``) over existing
        ``BTEDevice`` objects whose counters are views ``n_active[slots[p]]``
        and ``norm2[slots[p]]`` of the given arrays (the serving loop keeps
        problems resident and rebuilds only this view when its set changes)."""
        if not torch.cuda.is_available():
            raise RuntimeError('openbte_amd.BTEBatch needs a ROCm GPU; there is no CPU fallback')
        self.dev = device or torch.device('cuda', torch.cuda.current_device())
        self.cfg = cfg or SweepConfig()
        if devices is None:
            P = len(ops)
            self.n_active = torch.zeros(P, dtype=torch.int32, device=self.dev)
            self.norm2 = torch.zeros(P, dtype=torch.float64, device=self.dev)
            self.slots = np.arange(P)
            self.devs: List[BTEDevice] = []
            for i, op in enumerate(ops):
                sc = scheds[i] if scheds is not None else build_schedule(
                    op, max_width=batch_max_width(self.cfg.mc, self.cfg.block))
                d = BTEDevice(op, sched=sc, cfg=SweepConfig(**vars(self.cfg)),
                              device=self.dev, n_active=self.n_active[i:i + 1], norm2=self.norm2[i:i + 1])
                self.devs.append(d)
            assemble = True
        else:
            self.devs = list(devices)
            P = len(self.devs)
            self.n_active, self.norm2 = n_active, norm2
            self.slots = np.asarray(slots if slots is not None else np.arange(P), dtype=np.int64)
            for d in self.devs:
                if (d.cfg.mc, d.cfg.block) != (self.cfg.mc, self.cfg.block):
                    raise ValueError('every problem of a batch needs the batch (mc, block)')
            assemble = False  # each device assembled its operator when it was built
        self.n_problems = P
        # a problem with fewer inflow slots carries zero coefficients in the
        # rest of its records, so the widest launch serves every problem
        self.kin = max(d.kin for d in self.devs)
        self.stream = torch.cuda.current_stream(self.dev).cuda_stream
        self.live = torch.ones(P, dtype=torch.int32, device=self.dev)
        self._build_tasks()
        self._build_structs(groups, assemble)
        self.ext_max = torch.zeros(1, dtype=torch.int32, device=self.dev)
        self.n_iter = np.zeros(P, dtype=np.int64)
        self.error = np.ones(P)
        self.norm_old = np.zeros(P)

    # -- task table --------------------------------------------------------
    def wave_pair_limit(self) -> int:
        """Largest ``lvl_pair_max`` of a problem whose tasks run one per wave
        (0: none): its window, a power of two, must fit block / 64 times in
        the 48 KiB window budget."""
        cfg = self.cfg
        if not (cfg.wave_tasks and cfg.window) or cfg.mc > 16 or cfg.block < 128:
            return 0
        waves = cfg.block // 64
        cap = 64
        while waves * ((2 * cap) * cfg.mc // 2 + 1) * 16 <= 48 * 1024:
            cap *= 2
        return cap if waves * (cap * cfg.mc // 2 + 1) * 16 <= 48 * 1024 else 0

    def _build_tasks(self):
        """Task table: workgroup tasks in 8 queues, then wave tasks (problems
        whose adjacent levels hold at most ``wave_pair_limit()`` cells) in 8
        more; each set balanced greedily by cost, longest first."""
        lim = self.wave_pair_limit()
        self.wave_problem = np.array([0 < d.lvl_pair_max <= lim for d in self.devs], dtype=bool)
        sets = ([], [])  # (cost, pid, l, nch) of workgroup tasks, of wave tasks
        for pid, d in enumerate(self.devs):
            lev = d.sched.n_levels
            wave = bool(self.wave_problem[pid])
            for l in range(d.op.n_dirs):
                # passes per level: cells / (cells of one pass), at least one
                per_pass = 32 if wave else d.cfg.block
                cost = float(lev[l]) * max(1.0, d.op.n / (float(lev[l]) * per_pass)) * d.n_chunks
                sets[wave].append((cost, pid, l, d.n_chunks))
        tp, tl = [], []
        offs, loads = [], []
        for items in sets:
            items.sort(key=lambda t: -t[0])
            load = np.zeros(N_QUEUES)
            queues = [[] for _ in range(N_QUEUES)]
            for cost, pid, l, nch in items:
                q = int(np.argmin(load))
                load[q] += cost
                queues[q].append((cost, pid, l, nch))
            off = [len(tp)]
            for q in queues:
                for cost, pid, l, nch in q:
                    for c in range(nch):
                        tp.append(pid)
                        tl.append(l * nch + c)
                off.append(len(tp))
            offs.append(off)
            loads.append(load)
        self.task_pid = torch.as_tensor(np.asarray(tp, np.int32)).to(self.dev)
        self.task_local = torch.as_tensor(np.asarray(tl, np.int32)).to(self.dev)
        self.queue_off = torch.as_tensor(np.asarray(offs[0], np.int32)).to(self.dev)
        self.queue_cnt = torch.zeros(N_QUEUES, dtype=torch.int32, device=self.dev)
        self.n_tasks = len(tp)
        self.n_block_tasks = offs[0][-1]
        self.n_wave_tasks = self.n_tasks - self.n_block_tasks
        if self.n_wave_tasks:
            self.wave_off = torch.as_tensor(np.asarray(offs[1], np.int32)).to(self.dev)
            self.wave_cnt = torch.zeros(N_QUEUES, dtype=torch.int32, device=self.dev)
            self.wave_pair_max = max(d.lvl_pair_max for d, w in zip(self.devs, self.wave_problem) if w)
        else:
            self.wave_off = self.wave_cnt = None
            self.wave_pair_max = 0
        self.queue_load = loads[0]
        self.wave_load = loads[1]

    def _build_structs(self, groups, assemble: bool = True):
        devs = self.devs
        P = self.n_problems
        for d in devs:
            # fixed state pointers: DeltaT/TB in slot 0, the reduction writes slot 1
            d.plan.delta_t = d.buf['dt'][0].data_ptr()
            d.plan.tb = d.buf['tb'][0].data_ptr()
        self.plans_t = _struct_tensor([d.plan for d in devs], self.dev)
        self.gcrs_t = _struct_tensor([d.gcr for d in devs], self.dev)
        io = []
        for mode in range(3):
            for d in devs:
                g = d.gbuf
                s = _lib.SweepIO()
                if mode == 1:
                    s.y, s.x = g['w'].data_ptr(), g['x'].data_ptr()
                    s.active, s.lane_active = g['active'].data_ptr(), g['lane_active'].data_ptr()
                else:
                    s.y, s.x = g['y'].data_ptr(), g['x'].data_ptr()
                io.append(s)
        self.io_t = _struct_tensor(io, self.dev)
        rio = []
        for d in devs:
            r = _lib.ReduceIO()
            r.dt_new = d.buf['dt'][1].data_ptr()
            r.j_new = d.buf['j'].data_ptr()
            r.tb_new = d.buf['tb'][1].data_ptr()
            r.norm2 = d.norm2.data_ptr()
            rio.append(r)
        self.rio_t = _struct_tensor(rio, self.dev)
        B = self.b = _lib.Batch()
        B.n_problems, B.n_tasks = P, self.n_tasks
        B.mc, B.block, B.kin = self.cfg.mc, self.cfg.block, self.kin
        props = torch.cuda.get_device_properties(self.dev)
        cus = getattr(props, 'multi_processor_count', 256)
        waves = self.cfg.block // 64
        work = self.n_block_tasks + (self.n_wave_tasks + waves - 1) // waves
        B.groups = int(groups or min(work, cus * max(1, 1024 // self.cfg.block)))
        B.max_elems = max(d.op.n * d.op.n_dirs for d in devs)
        B.max_dirs = max(d.op.n_dirs for d in devs)
        B.max_cells = max(d.op.n for d in devs)
        B.max_walls = max(len(d.op.i_diffuse) for d in devs)
        B.lvl_pair_max = max(d.lvl_pair_max for d in devs)
        B.sweep_window = int(bool(self.cfg.window))
        B.part_sums = int(all(d.plan.part_chunk for d in devs))
        B.plans, B.gcrs, B.io, B.rio = (self.plans_t.data_ptr(), self.gcrs_t.data_ptr(), self.io_t.data_ptr(),
                                        self.rio_t.data_ptr())
        B.live = self.live.data_ptr()
        B.task_pid, B.task_local = self.task_pid.data_ptr(), self.task_local.data_ptr()
        B.queue_off, B.queue_cnt = self.queue_off.data_ptr(), self.queue_cnt.data_ptr()
        if self.wave_off is not None:
            B.wave_off, B.wave_cnt = self.wave_off.data_ptr(), self.wave_cnt.data_ptr()
            B.wave_pair_max = self.wave_pair_max
        if assemble:
            _lib.call('obte_batch_assemble', self.b, self.stream)

    def keep_solutions(self, on: bool = True) -> None:
        """``BTEDevice.keep_solutions`` for every problem of the batch."""
        for d in self.devs:
            d.keep_solutions(on)
        self.plans_t.copy_(_struct_tensor([d.plan for d in self.devs], self.dev))

    # -- iteration ---------------------------------------------------------
    def reset(self, delta_t0: Sequence[np.ndarray]) -> None:
        for d, t0 in zip(self.devs, delta_t0):
            x = torch.as_tensor(np.asarray(t0, dtype=np.float64)).to(self.dev)
            d.buf['dt'][0].copy_(x)
            if len(d.op.i_diffuse):
                d.buf['tb'][0][:len(d.op.i_diffuse)].copy_(x[torch.as_tensor(d.op.i_diffuse).to(self.dev)])
        self.live.fill_(1)
        self.n_iter[:] = 0
        self.error[:] = 1.0
        self.norm_old = np.array([float(np.linalg.norm(t)) for t in delta_t0])
        _lib.call('obte_batch_gcr_reset', self.b, 1, self.stream)

    # launch sequences (always on the caller's current stream, so they can be captured)
    def _head(self, rounds: int, log: bool = False):
        b, st = self.b, torch.cuda.current_stream(self.dev).cuda_stream
        _lib.call('obte_batch_source', b, st)
        _lib.call('obte_batch_sweep', b, 0, st)
        _lib.call('obte_batch_gcr', b, 0, st)
        if log:
            _lib.call('obte_batch_log_active', b, self.round_log[0:1].data_ptr(), st)
        for i in range(rounds):
            self._round()
            if log:
                _lib.call('obte_batch_log_active', b, self.round_log[i + 1:i + 2].data_ptr(), st)

    def _round(self):
        b, st = self.b, torch.cuda.current_stream(self.dev).cuda_stream
        _lib.call('obte_batch_gcr', b, 1, st)
        _lib.call('obte_batch_sweep', b, 1, st)
        _lib.call('obte_batch_gcr', b, 2, st)

    def _tail(self):
        b, st = self.b, torch.cuda.current_stream(self.dev).cuda_stream
        _lib.call('obte_batch_sweep', b, 2, st)
        _lib.call('obte_batch_reduce', b, st)

    def _fused(self, max_extend: int):
        """The whole outer iteration with the extension loop in one launch
        (``obte_batch_extend``)."""
        b, st = self.b, torch.cuda.current_stream(self.dev).cuda_stream
        _lib.call('obte_batch_source', b, st)
        _lib.call('obte_batch_sweep', b, 0, st)
        _lib.call('obte_batch_gcr', b, 0, st)
        _lib.call('obte_batch_extend', b, max_extend, self.ext_max.data_ptr(), st)
        self._tail()

    def _fused_done(self, max_extend: int) -> int:
        ext = int(self.ext_max.item())
        n_left = int(self.n_active.sum().item())
        if n_left > 0:
            warn_unconverged(n_left, max_extend)
        return ext

    def outer_iteration(self, max_extend: Optional[int] = None) -> int:
        """One outer iteration of every live problem; returns the extensions
        (the largest count of a task)."""
        max_extend = self.cfg.max_extend if max_extend is None else max_extend
        if self.cfg.fused_extend:
            if self.use_graphs:
                g = self._graphs.get(('fused', max_extend))
                if g is None:
                    g = self._graphs[('fused', max_extend)] = self._capture(self._fused, max_extend)
                g.replay()
            else:
                self._fused(max_extend)
            return self._fused_done(max_extend)
        if self.use_graphs:
            return self._outer_iteration_graph(max_extend)
        self._head(0)
        ext = 0
        while ext < max_extend and int(self.n_active.sum().item()) > 0:
            self._round()
            ext += 1
        if ext >= max_extend and int(self.n_active.sum().item()) > 0:
            warn_unconverged(int(self.n_active.sum().item()), max_extend)
        self._tail()
        return ext

    # -- hipGraph-captured outer iteration ---------------------------------
    use_graphs = False

    def enable_graphs(self, on: bool = True):
        """Replay the outer iteration from hipGraphs.  With the fused
        extension loop (``SweepConfig.fused_extend``) one graph holds the whole
        iteration.  Otherwise graph A = source, first sweep, projection and E
        extension rounds (rounds of converged tasks are no-ops inside the
        kernels); graph B = final sweep and reduction.  E follows the previous
        iteration's extension count; a residual left after graph A is finished
        by host-driven rounds before graph B, so the solve stays exact."""
        self.use_graphs = on
        self._graphs = {}
        self._tail_graph = None
        self._next_rounds = 8
        self.round_log = torch.zeros(self.cfg.max_extend + 2, dtype=torch.int32, device=self.dev)

    def _capture(self, fn, *args):
        # one capture at a time in the process: capture begins with a device
        # synchronisation, which is not allowed while another thread captures
        with _CAPTURE_LOCK:
            return self._capture_locked(fn, *args)

    def _capture_locked(self, fn, *args):
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.dev)
        s.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s, capture_error_mode='thread_local'):
                fn(*args)
        torch.cuda.current_stream(self.dev).wait_stream(s)
        return g

    def _outer_iteration_graph(self, max_extend: int) -> int:
        E = max(1, min(self._next_rounds, max_extend))
        if E not in self._graphs:
            self._graphs[E] = self._capture(self._head, E, True)
        if self._tail_graph is None:
            self._tail_graph = self._capture(self._tail)
        self._graphs[E].replay()
        log = self.round_log[:E + 1].cpu().numpy()
        ext = int((log[:E] > 0).sum())  # rounds that found active tasks
        extra = 0
        if log[E] > 0:
            while ext + extra < max_extend and int(self.n_active.sum().item()) > 0:
                self._round()
                extra += 1
            if ext + extra >= max_extend and int(self.n_active.sum().item()) > 0:
                warn_unconverged(int(self.n_active.sum().item()), max_extend)
        self._tail_graph.replay()
        self._next_rounds = max(1, ext + extra)
        return ext + extra

    def solve(self, delta_t0, bte_max_iter: int = 100, bte_rtol: float = 1e-4, on_iteration=None):
        """Iterate every problem to its own stop rule (openbte/bte.py:193-196,251-254)."""
        self.reset(delta_t0)
        live = np.ones(self.n_problems, bool)
        while live.any():
            self.outer_iteration()
            nrm = np.sqrt(self.norm2.cpu().numpy()[self.slots])
            for p in np.nonzero(live)[0]:
                self.n_iter[p] += 1
                with np.errstate(divide='ignore', invalid='ignore'):  # zero field: nan ends the loop
                    self.error[p] = abs(nrm[p] - self.norm_old[p]) / nrm[p]
                self.norm_old[p] = nrm[p]
                if not (self.error[p] > bte_rtol and self.n_iter[p] < bte_max_iter):
                    live[p] = False
            self.live.copy_(torch.as_tensor(live.astype(np.int32)))
            if on_iteration is not None:
                on_iteration(self, live)
        return self.n_iter.copy()

    def results(self, p: int) -> SolverResults:
        d = self.devs[p]
        op = d.op
        K = d.buf['kappa'].view(op.n_mfp, op.n_dirs).cpu().numpy().copy()
        S = d.buf['supp'].view(op.n_mfp, op.n_dirs).cpu().numpy().copy()
        DeltaT = d.buf['dt'][0].cpu().numpy().copy()
        J = d.buf['j'].view(op.n, op.dim).cpu().numpy().copy()
        if op.contact_kind == 0:
            K[:] = 0
            S[:] = 0
        kappa = float(K.sum())
        return SolverResults(kappa_eff=kappa,
                             variables={'Temperature_BTE': {'data': DeltaT, 'units': 'K'},
                                        'Flux_BTE': {'data': J * 1e9, 'units': 'W/m/m'}},
                             aux={'suppression': S, 'kappa_eff': kappa, 'n_iter': int(self.n_iter[p])})


def BTE_RTA_batch(problems, bte_max_iter: int = 100, bte_rtol: float = 1e-4,
                  config: Optional[SweepConfig] = None, autotune: bool = True,
                  graphs: bool = True) -> List[SolverResults]:
    """Solve several problems together.  ``problems`` is a sequence of dicts
    with the ``BTE_RTA`` arguments (geo, mat, bcs, fourier,
    effective_thermal_conductivity, heat_source).  Without ``config`` the
    launch shape is picked by ``tune.tune`` on these problems (memoised per
    problem signature; ``autotune=False``: the single-solve default, which
    makes every result bitwise that of ``BTE_RTA`` alone); ``graphs`` replays
    the outer iterations from hipGraphs."""
    ops, t0 = [], []
    for pr in problems:
        geo = pr['geo']
        ops.append(build_host_operator(geo, pr['mat'], pr['bcs'], pr.get('effective_thermal_conductivity'),
                                       pr.get('heat_source')))
        f = pr.get('fourier')
        t0.append(np.zeros(geo.n_elems) if f is None else
                  np.asarray(f.variables['Temperature_Fourier']['data'], np.float64))
    if config is None:
        if autotune:
            from . import tune
            config = tune.tune(ops)
        else:
            config = default_config(ops)
    batch = BTEBatch(ops, cfg=config)
    if graphs:
        batch.enable_graphs(True)
    batch.solve(t0, bte_max_iter=bte_max_iter, bte_rtol=bte_rtol)
    return [batch.results(p) for p in range(len(problems))]
