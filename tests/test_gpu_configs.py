"""Whole benchmark items against the oracle: c1 (2^n - 25, p from 1e9, 2^26 integers, classes
off; 2224 relations) and c2 (10^n + 111, p from 1e12, 2^28 integers, classes on; 12 relations),
run as the benchmark runs them (2^26-integer chunks, Worker.hpp:152-157) under class policies 1 and
2.  The relation counts are the oracle's, which tests/test_planted_cpu.py and tests/test_oracle.py
pin independently."""
import pytest

pytestmark = pytest.mark.gpu

CONFIGS = {
    "c1": ((2, -25, 10**9, 10**9 + 2**26, 3_600_000, 5_000_000, 2, 1), False, 2224),
    "c2": ((10, 111, 10**12, 10**12 + 2**28, 0, 10_000_000, 6, 4), True, 12),
}
_want = {}


def oracle_relations(oracle, name):
    if name not in _want:
        it = CONFIGS[name][0]
        _want[name], _ = oracle.process_range(oracle.make_item(*it), it[2], it[3], 1 << 22, 16)
    return _want[name]


@pytest.mark.parametrize("policy", [1, 2])
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_benchmark_item_matches_oracle(gpu, mps, oracle, name, policy):
    it, classes, count = CONFIGS[name]
    print(f"{name} policy {policy}: {it}", flush=True)
    gpu.set_classes(classes)
    gpu.set_class_policy(policy)
    try:
        rels, st = gpu.run_item(mps.work_item(*it), chunk_ints=1 << 26)
    finally:
        gpu.set_classes(True)
        gpu.set_class_policy(1)
    want = oracle_relations(oracle, name)
    assert len(want) == count
    assert st.declined == 0
    assert rels == want
