"""The batched / served engine (include/mps_batch.h): relations equal the oracle's CPU path item
by item, with and without hipGraph replay, alone and inside a mix of concurrently admitted items;
admission respects the active-item limit and the memory budget; ranges outside (3, 2^62) are
declined as the plain worker declines them."""
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "mass-prime-siever_amd"))
import mix  # noqa: E402

# README item families and the far end of the 62-bit range, each cut into several 2^20-integer
# chunks so that every slot replays its graphs many times
ITEMS = [
    (2, -25, 10**9, 10**9 + 2**22, 3_600_000, 5_000_000, 2, 1),
    (2, -25, 10**9, 10**9 + 2**21, 3_600_000, 5_000_000, 6, 1),
    (10, 111, 10**12, 10**12 + 2**22, 0, 10_000_000, 6, 4),
    (10, 111, 10**12, 10**12 + 2**21, 0, 10_000_000, 1, 0),
    (2, -25, 95 * 10**12, 95 * 10**12 + 2**21, 3_600_000, 5_000_000, 2, 1),
    (2, -25, 10**6, 10**6 + 2**21, 3_600_000, 5_000_000, 2, 1),
    (3, -2, 4 * 10**18, 4 * 10**18 + 2**20, 0, 10_000_000, 1, 0),
    (3, -2, (1 << 50) - 2**20, (1 << 50) + 2**20, 0, 1_000_000, 1, 0),  # crosses the FP64 field limit
]

PASS_ODDS = 1 << 19


def oracle_rels(oracle, it):
    b, c, P0, P1, N0, N1, mod, res = it
    want, _ = oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, mod, res), P0, P1, 1 << 20, 8)
    return want


@pytest.fixture(scope="module")
def want(oracle):
    return {it: oracle_rels(oracle, it) for it in ITEMS}


def mix_items(n, width):
    """The first n items of the seed-0 serving mix with their P ranges cut to width integers."""
    out = []
    for d in mix.serving_mix(0, n):
        out.append((d["b"], d["c"], d["P0"], d["P0"] + width, d["N0"], d["N1"], d["mod"], d["res"]))
    return out


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_matches_oracle(gpu, mps, want, graphs):
    print(f"batch graphs={graphs}", flush=True)
    eng = mps.Batch(gpu, slots=4, pass_odds=PASS_ODDS, use_graphs=graphs)
    try:
        ids = [eng.submit(mps.work_item(*it), tag=k) for k, it in enumerate(ITEMS)]
        eng.wait_all()
        done = {d["id"]: d for d in eng.poll()}
        assert sorted(done) == sorted(ids)
        for i, it in zip(ids, ITEMS):
            assert done[i]["status"] == mps.MPS_OK, (it, eng.last_error())
            assert eng.relations(i) == want[it], it
            assert done[i]["relations"] == len(want[it])
        st = eng.stats()
        chunks = sum(-(-(it[3] - it[2]) // (2 * PASS_ODDS)) for it in ITEMS)
        assert st["passes"] == chunks
        if graphs:
            assert st["graph_launches"] == chunks
            assert 0 < st["graphs_built"] < chunks  # graphs are replayed, not rebuilt per pass
        else:
            assert st["graph_launches"] == 0
    finally:
        eng.close()


def test_each_result_same_alone_and_in_mix(gpu, mps, oracle):
    """Continuous batching: an item's relations do not depend on what shares the device with it."""
    items = mix_items(12, 2**22)
    print("mix alone vs together", flush=True)
    alone = {}
    for it in items[:4]:
        eng = mps.Batch(gpu, slots=4, pass_odds=PASS_ODDS)
        try:
            i = eng.submit(mps.work_item(*it))
            eng.wait_all()
            alone[it] = eng.relations(i)
        finally:
            eng.close()
    eng = mps.Batch(gpu, slots=4, pass_odds=PASS_ODDS, policy=1)
    try:
        ids = [eng.submit(mps.work_item(*it)) for it in items]
        eng.wait_all()
        together = {it: eng.relations(i) for i, it in zip(ids, items)}
        statuses = {d["id"]: d["status"] for d in eng.poll()}
        assert all(s == mps.MPS_OK for s in statuses.values())
    finally:
        eng.close()
    for it in items[:4]:
        assert together[it] == alone[it], it
    for it in items[:6]:
        assert together[it] == oracle_rels(oracle, it), it
    for it in items[6:]:  # the plain per-chunk worker on the rest
        rels, _ = gpu.run_item(mps.work_item(*it), chunk_ints=1 << 21)
        assert together[it] == rels, it


def overlap_max(done):
    """Most items in progress (admitted, not finished) at any admission instant."""
    return max(sum(1 for e in done if e["t_admit"] <= d["t_admit"] < e["t_done"]) for d in done)


@pytest.mark.parametrize("limit", ["max_active", "mem_budget"])
def test_admission_limits(gpu, mps, want, limit):
    kw = {"max_active": 2} if limit == "max_active" else {"mem_budget": 2 * 49 * 32 * 8}
    print(f"admission {kw}", flush=True)
    eng = mps.Batch(gpu, slots=4, pass_odds=PASS_ODDS, **kw)
    try:
        ids = [eng.submit(mps.work_item(*it)) for it in ITEMS[:5]]
        eng.wait_all()
        done = eng.poll()
        assert len(done) == 5
        assert overlap_max(done) <= 2
        assert all(d["t_submit"] <= d["t_admit"] <= d["t_done"] for d in done)
        for i, it in zip(ids, ITEMS[:5]):
            assert eng.relations(i) == want[it]
    finally:
        eng.close()


def test_declined_ranges(gpu, mps):
    """gpu_range_ok (Worker.hpp:209): chunks outside (3, 2^62) are declined, the rest still run."""
    eng = mps.Batch(gpu, slots=2, pass_odds=PASS_ODDS)
    try:
        a = eng.submit(mps.work_item(2, -25, 2, 2**20, 3_600_000, 5_000_000, 2, 1))
        b = eng.submit(mps.work_item(3, -2, (1 << 62) - 2**20, (1 << 62) + 2**20, 0, 1000, 1, 0))
        c = eng.submit(mps.work_item(2, -25, 10**9, 10**9, 3_600_000, 5_000_000, 2, 1))  # empty range
        eng.wait_all()
        st = {d["id"]: d["status"] for d in eng.poll()}
        assert st[a] == mps.MPS_DECLINED
        assert st[b] == mps.MPS_DECLINED
        assert st[c] == mps.MPS_OK and eng.relations(c) == []
    finally:
        eng.close()


def test_batch_order_mode_matches_oracle(gpu, mps, want):
    """The batch engine under class policy 2 (order mode, its own setup kernel variant and graphs)."""
    gpu.set_class_policy(2)
    eng = mps.Batch(gpu, slots=4, pass_odds=PASS_ODDS)
    try:
        ids = [eng.submit(mps.work_item(*it)) for it in ITEMS]
        eng.wait_all()
        assert all(d["status"] == mps.MPS_OK for d in eng.poll())
        for i, it in zip(ids, ITEMS):
            assert eng.relations(i) == want[it], it
    finally:
        eng.close()
        gpu.set_class_policy(1)


@pytest.mark.parametrize("policy", [1, 2])
def test_batch_edge_items(gpu, mps, oracle, policy):
    """The degenerate and boundary items of the plain worker's edge test (b = 0 and 1, a negative
    base, N0 < 0, cls_res >= cls_mod, one in-class exponent, cm = 2 * 65537, an empty exponent
    range) and c = INT64_MIN / INT64_MAX, all in one batch: every item finishes with the oracle's
    relations and none is declined."""
    from test_gpu_relations import EDGE_ITEMS, EXTREME_C
    items = EDGE_ITEMS + EXTREME_C
    print(f"batch edge items, class policy {policy}", flush=True)
    gpu.set_class_policy(policy)
    eng = mps.Batch(gpu, slots=4, pass_odds=PASS_ODDS)
    try:
        ids = [eng.submit(mps.work_item(*it), tag=k) for k, it in enumerate(items)]
        eng.wait_all()
        done = {d["id"]: d for d in eng.poll()}
        for i, it in zip(ids, items):
            assert done[i]["status"] == mps.MPS_OK, (it, eng.last_error())
            want, _ = oracle.process_range(oracle.make_item(*it), it[2], it[3], 1 << 16, 8)
            assert eng.relations(i) == want, it
    finally:
        eng.close()
        gpu.set_class_policy(1)


def test_batch_planted_items(gpu, mps, oracle):
    """Every item of tests/golden/planted.json (relations planted at every magnitude, in both fields,
    small and large b and c) through the batch engine in order mode (class policy 2), submitted
    at once: each item's relations equal the oracle's over its P window, and a planted relation
    is among them."""
    import json
    import os
    planted = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "planted.json")))["items"]
    gpu.set_class_policy(2)
    eng = mps.Batch(gpu, slots=6, pass_odds=PASS_ODDS)
    try:
        print(f"batch: {len(planted)} planted items", flush=True)
        ids = [eng.submit(mps.work_item(*it["item"])) for it in planted]
        eng.wait_all()
        assert all(d["status"] == mps.MPS_OK for d in eng.poll())
        for i, it in zip(ids, planted):
            b, c, P0, P1, N0, N1, mod, res = it["item"]
            want, _ = oracle.process_range(oracle.make_item(*it["item"]), P0, P1, 1 << 17, 16)
            got = eng.relations(i)
            assert got == want, it["tag"]
            if it["relation"] is not None:
                assert tuple(it["relation"]) in got, it["tag"]
    finally:
        eng.close()
        gpu.set_class_policy(1)
