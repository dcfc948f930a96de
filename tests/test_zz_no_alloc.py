"""The allocation-free entry point mps_gpu_sieve_relations_noalloc (include/mps_gpu.h) against
mps_gpu_sieve_relations on the same ranges, through the C-ABI: two calls in a row on each item, the
second on a smaller range (the scratch array keeps the first call's contents, so the second shows
that nothing stale leaks through), at p ~ 1e12 (FP64 field) and ~4e18 (Montgomery field), with a
planted relation at 4e18 and one item with a relation at every prime (records past the staging
area).  Ranges above MPS_NOALLOC_MAX_INTS are refused before anything runs."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PLANTED = {it["tag"]: it for it in json.load(open(os.path.join(HERE, "golden", "planted.json")))["items"]}


def _items():
    c4 = PLANTED["c4/n5000000"]
    every = PLANTED["c2/n4"]
    return [
        ("c2", (10, 111, 0, 10_000_000, 6, 4), 10**12 + 12345),
        ("c1", (2, -25, 3_600_000, 5_000_000, 2, 1), 10**9 + 77),
        ("c4-planted", tuple(c4["item"][:2] + c4["item"][4:]), c4["item"][2]),
        ("every-prime", tuple(every["item"][:2] + every["item"][4:]), every["item"][2]),
    ]


@pytest.mark.parametrize("name,it,lo", _items(), ids=[x[0] for x in _items()])
def test_noalloc_matches_allocating_path(gpu, mps, name, it, lo):
    b, c, N0, N1, mod, res = it
    item = mps.work_item(b, c, lo, lo + mps.NOALLOC_MAX_INTS, N0, N1, mod, res)
    for width in (mps.NOALLOC_MAX_INTS, mps.NOALLOC_MAX_INTS // 2 - 1000):
        hi = lo + width
        print(f"no-alloc {name} [{lo}, {hi})", flush=True)
        st, got, n = gpu.sieve_relations_noalloc(lo, hi, item, cap=1 << 18)
        assert st == mps.MPS_OK, gpu.last_error()
        assert n == len(got)
        want = gpu.sieve_relations(lo, hi, item)
        assert want is not None
        assert sorted(got) == sorted(want)
        if name == "c4-planted" and width == mps.NOALLOC_MAX_INTS:
            rel = PLANTED["c4/n5000000"]["relation"]
            assert (rel[0], rel[1]) in got
        if name == "every-prime":
            assert len(got) > 4096
    # capacity: too small an output reports the count needed
    st, got, n = gpu.sieve_relations_noalloc(lo, lo + mps.NOALLOC_MAX_INTS // 2 - 1000, item, cap=1)
    assert (st == mps.MPS_CAPACITY and n > 1) or (st == mps.MPS_OK and n <= 1)


def test_noalloc_refuses_a_range_past_the_limit(gpu, mps):
    lo = 10**12
    item = mps.work_item(10, 111, lo, lo + 2**21, 0, 10_000_000, 6, 4)
    st, got, n = gpu.sieve_relations_noalloc(lo, lo + mps.NOALLOC_MAX_INTS + 2, item)
    assert st == mps.MPS_ERROR and n == 0
    assert "2^20" in gpu.last_error()
    # the handle still works afterwards
    st, got, n = gpu.sieve_relations_noalloc(lo, lo + 2**16, item)
    assert st == mps.MPS_OK
