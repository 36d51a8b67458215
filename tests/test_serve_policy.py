"""Host-side policy of the serving loop (no GPU): which requests end by
cancellation or deadline, and first-come-first-served admission by device
footprint under a byte budget and a slot count."""
import pytest

from openbte_amd.serve import Request, RequestCancelled, RequestTimeout, ending_error, plan_admission


def _req(t_submit, timeout_s=None, cancelled=False):
    r = Request(geo=None, mat=None, bcs=None, timeout_s=timeout_s)
    r.id, r.t_submit, r.cancelled = 7, t_submit, cancelled
    return r


def test_ending_error():
    assert ending_error(_req(10.0), now=1e9) is None                      # no deadline
    assert ending_error(_req(10.0, timeout_s=5.0), now=14.9) is None      # before the deadline
    assert isinstance(ending_error(_req(10.0, timeout_s=5.0), now=15.0), RequestTimeout)
    assert isinstance(ending_error(_req(10.0, timeout_s=0.0), now=10.0), RequestTimeout)
    # cancellation wins over a passed deadline
    assert isinstance(ending_error(_req(10.0, timeout_s=1.0, cancelled=True), now=20.0), RequestCancelled)


@pytest.mark.parametrize('sizes,slots,used,budget,expect', [
    ([3, 4, 2], 8, 0, 10, ([0, 1, 2], [], [])),
    ([6, 5, 3], 8, 0, 10, ([0, 2], [], [1])),       # a request that does not fit yet is skipped, not blocking
    ([11, 2], 8, 0, 10, ([1], [0], [])),            # larger than the whole budget: never admitted
    ([1, 1, 1], 2, 0, 10, ([0, 1], [], [2])),       # slot-bound
    ([4, 4], 8, 5, 10, ([0], [], [1])),             # memory already in use counts
])
def test_plan_admission(sizes, slots, used, budget, expect):
    assert plan_admission(sizes, slots, used, budget) == expect
