"""Autotuner of the sweep launch shape.

The sweep kernel has two shape knobs: ``mc`` (MFP bins per task: the width of
a cell's row in the solution workspace, hence MC/2 lanes per cell) and
``block`` (threads per workgroup, i.e. cells in flight per level of one task).
Narrow levels (small meshes) favour small blocks and many resident tasks; wide
levels favour 256-thread blocks.  The tuner times full sweeps (mode 0, all
tasks, the dominant launch of an outer iteration) for each candidate on the
actual problems, with HIP events on the launch stream, and keeps the fastest
per cell*mode update.  Results are cached in memory and, if ``cache`` names a
JSON file, on disk, keyed by the problem signature.
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .bte import BTEDevice, SweepConfig

CANDIDATES: Tuple[Tuple[int, int], ...] = ((4, 64), (4, 128), (8, 64), (8, 128), (8, 256), (16, 128), (16, 256))
_memo: Dict[str, dict] = {}


def signature(ops: Sequence) -> str:
    """Problem signature: sizes, mode counts and inflow widths of every operator."""
    rows = sorted((int(o.n), int(o.dim), int(o.n_mfp), int(o.n_dirs), int(len(o.i_diffuse))) for o in ops)
    return hashlib.sha1(json.dumps(rows).encode()).hexdigest()[:16]


def _time_launch(fn, reps: int) -> float:
    st = torch.cuda.current_stream()
    fn()  # warm-up
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / reps


def tune(ops: Sequence, scheds: Optional[Sequence] = None, candidates: Iterable[Tuple[int, int]] = CANDIDATES,
         reps: int = 3, cache: Optional[str] = None, verbose: bool = False) -> SweepConfig:
    """Pick (mc, block) for solving ``ops`` together (one problem: the plain
    path; several: the batched path).  Returns a ``SweepConfig``."""
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
            cfg = SweepConfig(mc=mc, block=block, kmax=8)
            if len(ops) == 1:
                d = BTEDevice(ops[0], sched=None if scheds is None else scheds[0], cfg=cfg)
                d.reset(np.zeros(ops[0].n))
                _lib.call('obte_bte_source', d.plan, d.stream)
                fn = (lambda d=d: _lib.call('obte_bte_sweep', d.plan, 0, d.gbuf['y'].data_ptr(),
                                            d.gbuf['x'].data_ptr(), 0, 0, torch.cuda.current_stream().cuda_stream))
            else:
                b = BTEBatch(ops, cfg=cfg, scheds=scheds)
                b.reset([np.zeros(o.n) for o in ops])
                _lib.call('obte_batch_source', b.b, b.stream)
                fn = (lambda b=b: _lib.call('obte_batch_sweep', b.b, 0, torch.cuda.current_stream().cuda_stream))
            t = _time_launch(fn, reps)
            results.append({'mc': mc, 'block': block, 'sweep_s': t, 'updates_per_s': updates / t})
            if verbose:
                print(f'# tune mc={mc} block={block}: sweep {t * 1e3:.3f} ms, {updates / t:.3e} updates/s',
                      flush=True)
            del fn
            torch.cuda.empty_cache()
        best = min(results, key=lambda r: r['sweep_s'])
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
