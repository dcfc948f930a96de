"""The plain worker's two pass lanes under concurrent calls on one handle.

Two threads call mps_gpu_sieve_relations at once with different items (different b, c, class
modulus, field and table shape), so each call runs on its own lane with its own descriptor, size
table, buffers and stream; every result must equal the records of the same call run alone, and a
third thread (more calls than lanes) waits for a lane and gets its own result too.  Then two
different items run through mps_run_item concurrently (four feeders on two lanes), against the
oracle.  The records are compared as sets of (p, M) with M exact or left to the host."""
import threading

import pytest

pytestmark = pytest.mark.gpu

ITEMS = [
    (10, 111, 10**12, 10**12 + 2**24, 0, 10**7, 6, 4),  # c2's kind: order mode, FP64 field
    (2, -25, 10**9, 10**9 + 2**24, 3_600_000, 5_000_000, 2, 1),  # c1's kind: many relations
    (3, -2, 4 * 10**18, 4 * 10**18 + 2**23, 0, 10**7, 1, 0),  # c4's kind: Montgomery field
]


def records(gpu, mps, it):
    recs = gpu.sieve_relations(it[2], it[3], mps.work_item(*it))
    assert recs is not None, gpu.last_error()
    return sorted(recs)


def test_concurrent_calls_on_two_lanes(gpu, mps):
    alone = [records(gpu, mps, it) for it in ITEMS]
    assert len(alone[1]) > 0
    for rep in range(3):
        print(f"concurrent calls, repetition {rep}", flush=True)
        got = [None] * len(ITEMS)
        errors = []

        def run(k):
            try:
                got[k] = records(gpu, mps, ITEMS[k])
            except Exception as e:  # reported below, in the main thread
                errors.append(repr(e))

        threads = [threading.Thread(target=run, args=(k,)) for k in range(len(ITEMS))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors, errors
        for k in range(len(ITEMS)):
            assert got[k] == alone[k], f"item {k} differs when run beside the others"


def test_concurrent_items_through_run_item(gpu, mps, oracle):
    its = ITEMS[:2]
    want = [oracle.process_range(oracle.make_item(*it), it[2], it[3], 1 << 18, 8)[0] for it in its]
    got = [None, None]
    errors = []

    def run(k):
        try:
            rels, st = gpu.run_item(mps.work_item(*its[k]), chunk_ints=1 << 22)
            assert st.declined == 0
            got[k] = rels
        except Exception as e:
            errors.append(repr(e))

    print("two items through run_item at once", flush=True)
    threads = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert got == want
