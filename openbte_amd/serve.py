"""Continuous-batching BTE server.

Requests (a geometry, material and boundary conditions each, as for
``BTE_RTA``, openbte/bte.py:19-27) arrive at any time.  Host preparation (the
operator of openbte/bte.py:47-158, the sweep schedule and, unless the request
brings one, the Fourier initial guess of openbte/fourier.py) runs in worker
threads; the Fourier solve uses its own stream, so it overlaps the device
work of the running problems.

The device loop works in rounds.  Each round

1. admits prepared requests while a slot is free and their device footprint
   (``bte.device_bytes``) fits in the HBM budget; an admitted problem keeps its
   mesh, operator, schedule, Krylov basis and per-mode state resident until it
   finishes;
2. runs one outer iteration (openbte/bte.py:193-256) of every admitted problem
   in shared launches (``BTEBatch`` over the resident devices: one persistent
   sweep launch covers every problem's (direction, MFP chunk) tasks);
3. applies each problem's own stop rule and retires the finished ones, whose
   memory and slot go to the next requests.

A request may carry a deadline (``timeout_s`` after submission) and may be
cancelled (``BTEServer.cancel``) at any time.  Both take effect at the next
round: the request ends with ``RequestTimeout`` or ``RequestCancelled`` in
``error`` and no result, and its slot and memory are released; the problems
that share the device with it are untouched (each problem's state and
reductions are its own).  The reference has no such path: a worker that dies
or hangs blocks its barrier forever (openbte/objects.py:21,88-101).

A problem therefore follows exactly the iteration it would follow alone, and
every kernel treats problems independently (fixed-order reductions, per-problem
counters), so its result does not depend on what else is in the batch.
"""
from __future__ import annotations

import contextlib
import heapq
import itertools
import os
import threading
import time
from concurrent.futures import FIRST_COMPLETED, CancelledError, ThreadPoolExecutor, wait
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

import numpy as np
import torch

from . import _lib
from .batch import BTEBatch
from .bte import BTEDevice, SweepConfig, device_bytes
from .fourier import Fourier
from .objects import SolverResults
from .operator import batch_max_width, build_host_operator, build_schedule


@dataclass
class Request:
    """One BTE solve.  ``fourier`` is a Fourier ``SolverResults`` to start
    from, ``'auto'`` to compute it, or ``None`` to start from DeltaT = 0
    (openbte/bte.py:38-43)."""
    geo: Any
    mat: Any
    bcs: Any
    effective_thermal_conductivity: Any = None
    heat_source: Any = None
    fourier: Any = 'auto'
    bte_max_iter: int = 100
    bte_rtol: float = 1e-4
    tag: Any = None
    timeout_s: Optional[float] = None   # deadline after submission (None: none)
    # filled in by the server
    id: int = -1
    t_submit: float = 0.0
    t_ready: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    bytes: int = 0
    updates: int = 0
    result: Optional[SolverResults] = None
    error: Optional[BaseException] = None
    cancelled: bool = False

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class RequestTimeout(TimeoutError):
    """A request passed its ``timeout_s`` before its solve finished."""


class RequestCancelled(CancelledError):
    """A request was cancelled before its solve finished."""


def ending_error(req: Request, now: float) -> Optional[BaseException]:
    """The error that ends ``req`` now, if any: cancellation first, then its
    deadline (``timeout_s`` after submission; 0 expires at the next round)."""
    if req.cancelled:
        return RequestCancelled(f'request {req.id} cancelled')
    if req.timeout_s is not None and now - req.t_submit >= req.timeout_s:
        return RequestTimeout(f'request {req.id} passed its {req.timeout_s:g} s deadline')
    return None


def plan_admission(sizes: List[int], n_free_slots: int, used: int, budget: int):
    """First come, first served admission by device footprint: returns the
    indices to admit now, those that can never fit (larger than the whole
    budget) and those that wait for memory or a slot."""
    admit, never, wait = [], [], []
    for i, b in enumerate(sizes):
        if b > budget:
            never.append(i)
        elif len(admit) < n_free_slots and used + b <= budget:
            admit.append(i)
            used += b
        else:
            wait.append(i)
    return admit, never, wait


@dataclass
class _Prepared:
    req: Request
    op: Any
    sched: Any
    t0: np.ndarray


@dataclass
class _Live:
    req: Request
    dev: BTEDevice
    slot: int
    n_iter: int = 0
    norm_old: float = 0.0
    error: float = 1.0


@dataclass
class ServerStats:
    submitted: int = 0
    completed: int = 0
    failed: int = 0          # ended without a result (preparation error, memory, timeout, cancellation)
    rounds: int = 0
    updates: int = 0
    busy_s: float = 0.0
    peak_bytes: int = 0
    peak_live: int = 0
    latencies: List[float] = field(default_factory=list)

    def percentile(self, q: float) -> float:
        return float(np.percentile(self.latencies, q)) if self.latencies else float('nan')


def host_cores() -> int:
    """Host cores this process may use (its affinity set)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # platforms without affinity sets
        return os.cpu_count() or 1


class BTEServer:
    """Continuous batching of BTE solves on one GPU.

    ``hbm_budget`` (bytes) bounds the summed device footprint of admitted
    problems (default: 85% of the free device memory at start); ``max_slots``
    bounds their number.  ``cfg`` fixes the launch shape (mc, block) shared by
    every problem.  With ``graphs`` the launches of a round are replayed from
    hipGraphs captured once per resident set (the set changes on admission
    and retirement).
    """

    def __init__(self, cfg: Optional[SweepConfig] = None, hbm_budget: Optional[int] = None, max_slots: int = 64,
                 prep_workers: int = 2, device=None, graphs: bool = True, fourier_concurrency: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError('openbte_amd.BTEServer needs a ROCm GPU; there is no CPU fallback')
        self.dev = device or torch.device('cuda', torch.cuda.current_device())
        self.cfg = cfg or SweepConfig()
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.dev)
            hbm_budget = int(0.85 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.n_active = torch.zeros(self.max_slots, dtype=torch.int32, device=self.dev)
        self.norm2 = torch.zeros(self.max_slots, dtype=torch.float64, device=self.dev)
        self.free_slots = list(range(self.max_slots))
        self.live: List[_Live] = []
        self.ready: List[_Prepared] = []
        self.preparing: List[Any] = []  # futures of (request, prepared or None)
        # requests waiting for a preparation thread, largest first
        self._prep_queue: List[tuple] = []
        self._prep_lock = threading.Lock()
        self.done: List[Request] = []
        self.used_bytes = 0
        self.stats = ServerStats()
        self._ids = itertools.count()
        # each preparation thread gets an equal share of the host cores for
        # the native scheduler's parallel loops (one thread per core in each
        # of them oversubscribes the host: 8 threads, 16 cores, serve mix of
        # 24 requests: makespan 2.5 s with 16 threads each, 1.9-2.1 s with 2)
        workers = max(1, prep_workers)
        share = max(1, host_cores() // workers)
        self._pool = ThreadPoolExecutor(max_workers=workers, initializer=_lib.call,
                                        initargs=('obte_set_host_threads', share))
        # Fourier guesses running on the device at once (0: no limit)
        self._fourier_gate = threading.Semaphore(fourier_concurrency) if fourier_concurrency > 0 else None
        self._batch: Optional[BTEBatch] = None
        self._dirty = True
        self.graphs = bool(graphs)  # replay each round's launches from hipGraphs (per resident set)

    # -- intake --------------------------------------------------------------
    def submit(self, req: Request) -> Request:
        """Queue a request (non-blocking) for host preparation.  A free
        preparation thread takes the largest waiting request first (cells x
        modes): the requests that need the most rounds are admitted first,
        which shortens the makespan of a burst of requests."""
        req.id = next(self._ids)
        req.t_submit = time.perf_counter()
        self.stats.submitted += 1
        try:
            cost = float(req.geo.n_elems) * float(np.asarray(req.mat.t_coeff).size)
        except Exception:  # a malformed request fails in its preparation, alone
            cost = 0.0
        with self._prep_lock:
            heapq.heappush(self._prep_queue, (-cost, req.id, req))
        self.preparing.append(self._pool.submit(self._prepare_next))
        return req

    def _prepare_next(self):
        with self._prep_lock:
            _, _, req = heapq.heappop(self._prep_queue)
        return req, self._prepare(req)

    def _prepare(self, req: Request) -> Optional[_Prepared]:
        try:
            op = build_host_operator(req.geo, req.mat, req.bcs, req.effective_thermal_conductivity,
                                     req.heat_source)
            sched = build_schedule(op, max_width=batch_max_width(self.cfg.mc, self.cfg.block))
            f = req.fourier
            if isinstance(f, str) and f == 'auto':
                gate = self._fourier_gate if self._fourier_gate is not None else contextlib.nullcontext()
                with gate:
                    stream = torch.cuda.Stream(device=self.dev)
                    with torch.cuda.stream(stream):
                        f = Fourier(req.geo, req.mat.thermal_conductivity, req.bcs,
                                    req.effective_thermal_conductivity, verbose=False)
                    stream.synchronize()
            t0 = (np.zeros(op.n) if f is None else
                  np.asarray(f.variables['Temperature_Fourier']['data'], np.float64))
            req.bytes = device_bytes(op, sched, self.cfg)
        except Exception as e:  # a request that cannot be prepared fails alone
            req.error = e
            return None
        req.t_ready = time.perf_counter()
        return _Prepared(req, op, sched, t0)

    def cancel(self, req: Request) -> None:
        """Cancel a submitted request; it ends at the next round (a request
        still being prepared ends when its preparation returns)."""
        req.cancelled = True

    def _collect_prepared(self) -> None:
        still = []
        for fut in self.preparing:
            if not fut.done():
                still.append(fut)
                continue
            req, p = fut.result()
            if p is None:
                self._finish_failed(req)
            else:
                self.ready.append(p)
        self.preparing = still

    def _reap(self) -> bool:
        """End the prepared and admitted requests that were cancelled or
        passed their deadline; returns True when a live problem ended."""
        now = time.perf_counter()
        keep = []
        for p in self.ready:
            err = ending_error(p.req, now)
            if err is None:
                keep.append(p)
            else:
                p.req.error = err
                self._finish_failed(p.req)
        self.ready = keep
        keep, ended = [], False
        for x in self.live:
            err = ending_error(x.req, now)
            if err is None:
                keep.append(x)
                continue
            x.req.error = err
            self._release(x)
            self._finish_failed(x.req)
            ended = True
        self.live = keep
        if ended:
            self._dirty = True
            self._batch = None  # drop the view's references to the released devices
        return ended

    def _finish_failed(self, req: Request) -> None:
        req.t_done = time.perf_counter()
        self.done.append(req)
        self.stats.completed += 1
        self.stats.failed += 1

    # -- admission -----------------------------------------------------------
    def _admit(self) -> None:
        # first come first served, skipping a request that does not fit yet
        admit, never, wait = plan_admission([p.req.bytes for p in self.ready], len(self.free_slots),
                                            self.used_bytes, self.hbm_budget)
        for i in never:
            req = self.ready[i].req
            req.error = MemoryError(f'request needs {req.bytes} bytes, budget is {self.hbm_budget}')
            self._finish_failed(req)
        keep = [self.ready[i] for i in wait]
        for i in admit:
            p = self.ready[i]
            req = p.req
            slot = self.free_slots.pop(0)
            dev = BTEDevice(p.op, sched=p.sched, cfg=SweepConfig(**vars(self.cfg)), device=self.dev,
                            n_active=self.n_active[slot:slot + 1], norm2=self.norm2[slot:slot + 1])
            dev.reset(p.t0)
            req.t_admit = time.perf_counter()
            self.used_bytes += req.bytes
            self.live.append(_Live(req, dev, slot, 0, float(np.linalg.norm(p.t0)), 1.0))
            self._dirty = True
        self.ready = keep
        self.stats.peak_bytes = max(self.stats.peak_bytes, self.used_bytes)
        self.stats.peak_live = max(self.stats.peak_live, len(self.live))

    # -- device rounds -------------------------------------------------------
    def _view(self) -> BTEBatch:
        if self._dirty or self._batch is None:
            self._batch = BTEBatch(cfg=self.cfg, device=self.dev, devices=[x.dev for x in self.live],
                                   n_active=self.n_active, norm2=self.norm2, slots=[x.slot for x in self.live])
            if self.graphs:
                self._batch.enable_graphs(True)
            self._dirty = False
        return self._batch

    def step(self) -> List[Request]:
        """Admit, run one outer iteration of every admitted problem, retire
        the finished ones; returns the requests completed in this round."""
        self._collect_prepared()
        self._reap()
        self._admit()
        if not self.live:
            if self.preparing:
                # nothing to run yet: wait for a preparation to finish rather
                # than return at once (a caller looping on step would hold the
                # interpreter lock the preparation threads need)
                wait(self.preparing, timeout=0.002, return_when=FIRST_COMPLETED)
            return []
        t = time.perf_counter()
        b = self._view()
        b.outer_iteration()
        nrm = np.sqrt(self.norm2.cpu().numpy())
        self.stats.rounds += 1
        finished, keep = [], []
        for x in self.live:
            x.n_iter += 1
            op = x.dev.op
            self.stats.updates += op.n * op.n_mfp * op.n_dirs
            x.req.updates += op.n * op.n_mfp * op.n_dirs
            n = nrm[x.slot]
            x.error = abs(n - x.norm_old) / n if n != 0 else 0.0
            x.norm_old = n
            # stop rule of openbte/bte.py:193-196,251-254
            if x.error > x.req.bte_rtol and x.n_iter < x.req.bte_max_iter:
                keep.append(x)
            else:
                finished.append(x)
        if finished:
            for x in finished:
                self._retire(x)
            self._dirty = True
            self._batch = None
        self.live = keep
        self.stats.busy_s += time.perf_counter() - t
        return [x.req for x in finished]

    def _retire(self, x: _Live) -> None:
        d = x.dev
        op = d.op
        K = d.buf['kappa'].view(op.n_mfp, op.n_dirs).cpu().numpy().copy()
        S = d.buf['supp'].view(op.n_mfp, op.n_dirs).cpu().numpy().copy()
        if op.contact_kind == 0:
            K[:] = 0
            S[:] = 0
        T = d.buf['dt'][d.cur].cpu().numpy().copy()
        J = d.buf['j'].view(op.n, op.dim).cpu().numpy().copy()
        kappa = float(K.sum())
        x.req.result = SolverResults(kappa_eff=kappa,
                                     variables={'Temperature_BTE': {'data': T, 'units': 'K'},
                                                'Flux_BTE': {'data': J * 1e9, 'units': 'W/m/m'}},
                                     aux={'suppression': S, 'kappa_eff': kappa, 'n_iter': x.n_iter})
        x.req.t_done = time.perf_counter()
        self._release(x)
        self.done.append(x.req)
        self.stats.completed += 1
        self.stats.latencies.append(x.req.latency)

    def _release(self, x: _Live) -> None:
        """Give a live problem's slot and device memory back."""
        self.n_active[x.slot:x.slot + 1].zero_()
        self.norm2[x.slot:x.slot + 1].zero_()
        self.free_slots.append(x.slot)
        self.free_slots.sort()
        self.used_bytes -= x.req.bytes
        x.dev = None

    def wait_prepared(self, poll_s: float = 0.001) -> None:
        """Block until every submitted request has finished host preparation."""
        while self.preparing:
            self._collect_prepared()
            if self.preparing:
                time.sleep(poll_s)

    def pending(self) -> int:
        return len(self.preparing) + len(self.ready) + len(self.live)

    def run_until_idle(self, poll_s: float = 0.001) -> List[Request]:
        out = []
        while self.pending():
            got = self.step()
            out.extend(got)
            if not self.live and not got:
                time.sleep(poll_s)
        return out

    def close(self) -> None:
        self._pool.shutdown(wait=True)


def summarize(reqs: List[Request]) -> Dict[str, float]:
    lat = np.array([r.latency for r in reqs if r.result is not None])
    return {'n': int(len(lat)), 'p50_s': float(np.percentile(lat, 50)) if len(lat) else float('nan'),
            'p99_s': float(np.percentile(lat, 99)) if len(lat) else float('nan'),
            'mean_s': float(lat.mean()) if len(lat) else float('nan')}
