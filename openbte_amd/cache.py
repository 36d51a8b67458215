"""Content-keyed, byte-bounded LRU caches of host preparation.

The level schedule of a mesh (one per direction set) and the Fourier
operator with its coarse level depend only on the mesh, the directions and
the conductivity tensor, not on the MFP bins, the drive or the iteration.
Solves that reuse a mesh (material or parameter sweeps, inverse-design steps,
served requests on one geometry) skip that work.  Keys are hashes of the
arrays' bytes, so a mutated array is a different key.
"""
from __future__ import annotations

import os
import threading
from collections import OrderedDict
from typing import Any, Callable, Iterable

import numpy as np

try:
    import xxhash

    def _hasher():
        return xxhash.xxh3_128()
except ImportError:  # pragma: no cover - the hash only keys a cache
    import hashlib

    def _hasher():
        return hashlib.blake2b(digest_size=16)


def content_key(*parts: Iterable[Any]) -> str:
    """Hash of numpy arrays (dtype, shape and bytes) and plain values."""
    h = _hasher()
    for p in parts:
        if isinstance(p, np.ndarray):
            a = np.ascontiguousarray(p)
            h.update(f'{a.dtype.str}{a.shape}'.encode())
            h.update(a.view(np.uint8).reshape(-1).data)
        else:
            h.update(repr(p).encode())
        h.update(b'|')
    return h.hexdigest()


def nbytes(obj: Any) -> int:
    """Bytes of the numpy arrays held by a NamedTuple / tuple / dict."""
    if isinstance(obj, np.ndarray):
        return int(obj.nbytes)
    if isinstance(obj, dict):
        return sum(nbytes(v) for v in obj.values())
    if isinstance(obj, (tuple, list)):
        return sum(nbytes(v) for v in obj)
    return 0


class ByteLRU:
    """Thread-safe LRU map bounded by the summed size of its values."""

    def __init__(self, max_bytes: int):
        self.max_bytes = int(max_bytes)
        self._d: "OrderedDict[str, tuple]" = OrderedDict()
        self._bytes = 0
        self._lock = threading.Lock()
        self.hits = 0
        self.misses = 0

    def get_or_build(self, key: str, build: Callable[[], Any]) -> Any:
        with self._lock:
            if key in self._d:
                self._d.move_to_end(key)
                self.hits += 1
                return self._d[key][0]
            self.misses += 1
        value = build()
        size = nbytes(value)
        if size <= self.max_bytes:
            with self._lock:
                if key not in self._d:
                    self._d[key] = (value, size)
                    self._bytes += size
                while self._bytes > self.max_bytes and self._d:
                    _, (_, sz) = self._d.popitem(last=False)
                    self._bytes -= sz
        return value

    def clear(self) -> None:
        with self._lock:
            self._d.clear()
            self._bytes = 0


SCHEDULES = ByteLRU(int(float(os.environ.get('OBTE_SCHEDULE_CACHE_BYTES', 2 * 2 ** 30))))
FOURIER = ByteLRU(int(float(os.environ.get('OBTE_FOURIER_CACHE_BYTES', 2 ** 30))))


def clear_all() -> None:
    """Forget every cached preparation (cold-start timings)."""
    SCHEDULES.clear()
    FOURIER.clear()
