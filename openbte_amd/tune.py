"""Autotuner of the sweep launch shape.

The sweep kernel has two shape knobs: ``mc`` (MFP bins per task: the width of
a cell's row in the solution workspace, hence MC/2 lanes per cell) and
``block`` (threads per workgroup, i.e. cells in flight per level of one task).
Narrow levels (small meshes) favour small blocks and many resident tasks; wide
levels favour 256-thread blocks; smaller chunks let transfer sweeps skip
converged bins at a finer grain.  The tuner times whole outer iterations (the
first ones of a solve from DeltaT = 0: source, transfer and final sweeps,
interface Krylov steps, mode sums) for each candidate on the actual problems,
with HIP events on the launch stream, and keeps the fastest.  Results are
cached in memory and, if ``cache`` names a JSON file, on disk, keyed by the
problem signature.
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .bte import BTEDevice, SweepConfig
from .operator import batch_max_width, build_schedule

CANDIDATES: Tuple[Tuple[int, int], ...] = ((4, 64), (4, 128), (8, 64), (8, 128), (8, 256), (16, 128), (16, 256))
_memo: Dict[str, dict] = {}


def signature(ops: Sequence) -> str:
    """Problem signature: sizes, mode counts and inflow widths of every operator."""
    rows = sorted((int(o.n), int(o.dim), int(o.n_mfp), int(o.n_dirs), int(len(o.i_diffuse))) for o in ops)
    return hashlib.sha1(json.dumps(rows).encode()).hexdigest()[:16]


def _time_iterations(obj, t0, n: int) -> float:
    """Seconds per outer iteration of ``obj`` (a BTEDevice or BTEBatch) over
    iterations 1..n from ``t0`` (after one untimed iteration that loads the
    kernels), with HIP events on the launch stream."""
    obj.reset(t0)
    obj.outer_iteration()
    obj.reset(t0)
    st = torch.cuda.current_stream()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(n):
        obj.outer_iteration()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / n


def tune(ops: Sequence, scheds: Optional[Sequence] = None, candidates: Iterable[Tuple[int, int]] = CANDIDATES,
         reps: int = 2, cache: Optional[str] = None, verbose: bool = False) -> SweepConfig:
    """Pick (mc, block) for solving ``ops`` together (one problem: the plain
    path; several: the batched path): the candidate with the shortest outer
    iterations 1..reps from DeltaT = 0.  Without ``scheds`` a batch candidate
    runs on schedules whose levels fit one pass of its batch kernel
    (``batch_max_width``).  Returns a ``SweepConfig``."""
    from .batch import BTEBatch
    key = signature(ops)
    rec = _memo.get(key)
    if rec is None and cache and os.path.exists(cache):
        with open(cache) as f:
            rec = json.load(f).get(key)
    if rec is None:
        updates = float(sum(o.n * o.n_mfp * o.n_dirs for o in ops))
        results: List[dict] = []
        for mc, block in candidates:
            cfg = SweepConfig(mc=mc, block=block)
            if len(ops) == 1:
                obj = BTEDevice(ops[0], sched=None if scheds is None else scheds[0], cfg=cfg)
                t0 = np.zeros(ops[0].n)
            else:
                sc = scheds if scheds is not None else [
                    build_schedule(o, max_width=batch_max_width(mc, block)) for o in ops]
                obj = BTEBatch(ops, cfg=cfg, scheds=sc)
                t0 = [np.zeros(o.n) for o in ops]
            t = _time_iterations(obj, t0, reps)
            results.append({'mc': mc, 'block': block, 'iter_s': t, 'updates_per_s': updates / t})
            if verbose:
                print(f'# tune mc={mc} block={block}: {t * 1e3:.3f} ms per outer iteration', flush=True)
            del obj
            torch.cuda.empty_cache()
        best = min(results, key=lambda r: r['iter_s'])
        rec = {'mc': best['mc'], 'block': best['block'], 'results': results}
        if cache:
            data = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    data = json.load(f)
            data[key] = rec
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, 'w') as f:
                json.dump(data, f, indent=1)
    _memo[key] = rec
    return SweepConfig(mc=rec['mc'], block=rec['block'])


def last_results(ops: Sequence) -> Optional[List[dict]]:
    """Timings behind the cached choice for ``ops`` (None if not tuned)."""
    rec = _memo.get(signature(ops))
    return None if rec is None else rec['results']
