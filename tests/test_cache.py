"""Content-keyed preparation caches: a schedule is reused for the same faces
and directions and rebuilt when any input array changes; the byte bound
evicts least recently used entries."""
import numpy as np

from openbte_amd.cache import ByteLRU, content_key


def test_content_key_tracks_bytes_dtype_and_shape():
    a = np.arange(12.0)
    assert content_key(a, 3) == content_key(a.copy(), 3)
    b = a.copy()
    b[5] += 1e-12
    assert content_key(a) != content_key(b)
    assert content_key(a) != content_key(a.astype(np.float32))
    assert content_key(a) != content_key(a.reshape(3, 4))
    assert content_key(a, 1) != content_key(a, 2)


def test_byte_lru():
    c = ByteLRU(max_bytes=3 * 800)
    built = []

    def make(k):
        built.append(k)
        return (np.zeros(100),)       # 800 bytes
    for k in ['a', 'b', 'c', 'a', 'd', 'b']:
        c.get_or_build(k, lambda k=k: make(k))
    # a, b, c built; a hit; d evicts b (least recent); b rebuilt
    assert built == ['a', 'b', 'c', 'd', 'b']
    assert c.hits == 1 and c.misses == 5
    big = c.get_or_build('huge', lambda: (np.zeros(1000),))   # larger than the bound: not kept
    assert big[0].size == 1000 and 'huge' not in c._d


def test_schedule_cache_reuse(small_case):
    from openbte_amd.cache import SCHEDULES
    from openbte_amd.operator import build_host_operator, build_schedule
    geo, mat, bcs, eff = small_case
    op = build_host_operator(geo, mat, bcs, eff)
    s1 = build_schedule(op)
    s2 = build_schedule(build_host_operator(geo, mat, bcs, eff))
    assert s1 is s2                                   # same faces and directions
    s3 = build_schedule(op, dirs=op.dirs[::-1].copy())
    assert s3 is not s1
    assert not np.array_equal(s3.order, s1.order)
    SCHEDULES.clear()
    assert build_schedule(op) is not s1
