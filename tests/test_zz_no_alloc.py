"""The allocation-free intersection entry point (asdm_intersect_k_noalloc, fixed per-GPU scratch)
against the allocating one (asdm_intersect_k) and numpy, through the C ABI: two calls in a row,
the second with a smaller input, then inputs beyond the scratch's limits, which are rejected
before any GPU work.  Runs kernels on the MI355X."""
from functools import reduce
import sys

import numpy as np
import pytest

import asdmatch_amd as A

pytestmark = pytest.mark.gpu


def _lists(rng, sizes, hi):
    return [np.unique(rng.integers(0, hi, size=n, dtype=np.uint64)).astype(np.uint32) for n in sizes]


def test_noalloc_intersection_matches_allocating_entry_point():
    ctx = A.Context(0)
    rng = np.random.default_rng(9)
    big = _lists(rng, (400_000, 700_000, 600_000), 1 << 21)        # about 1.6 Mi elements staged
    small = _lists(rng, (300, 2_000, 800, 1_500, 5_000), 6_000)    # stale data of the first call remains
    disjoint = [np.arange(0, 100, 2, dtype=np.uint32), np.arange(1, 100, 2, dtype=np.uint32)]
    for name, lists in (("big", big), ("small", small), ("disjoint", disjoint)):
        print(f"no-alloc intersection: {name}, {sum(len(a) for a in lists)} elements", flush=True, file=sys.stderr)
        want = reduce(np.intersect1d, lists)
        assert np.array_equal(ctx.intersect(lists), want), name
        got = ctx.intersect_noalloc(lists)
        assert np.array_equal(got, want), name
    # limits: more than 64 lists, more than 2^21 elements -> ASDM_E_BUDGET (-6), no GPU work
    with pytest.raises(A.AsdmError) as e:
        ctx.intersect_noalloc([np.arange(10, dtype=np.uint32)] * 65)
    assert e.value.code == -6
    with pytest.raises(A.AsdmError) as e:
        ctx.intersect_noalloc([np.arange(1 << 21, dtype=np.uint32), np.arange(10, dtype=np.uint32)])
    assert e.value.code == -6
    # still correct after the rejections
    assert np.array_equal(ctx.intersect_noalloc(small), reduce(np.intersect1d, small))
    ctx.close()
