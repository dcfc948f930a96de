"""End-to-end parity: relations of the GPU worker (device sieve + filter + BSGS, exact host
re-check) equal the oracle's CPU path (segmented sieve + run_bsgs on every prime) on the same
items, compared as sorted (p, M, q), as the reference's own self-test compares them
(kaggle/kaggle_run.py:64-95)."""
import pytest

pytestmark = pytest.mark.gpu

ITEMS = [
    # README item families (README.md:33-36): 2^n-25 with classes 2/1, 6/1 and none; 10^n+111 with
    # 6/4 and none; p from 10^6 where the sieve primes reach past sqrt(p).
    (2, -25, 10**9, 10**9 + 2**22, 3_600_000, 5_000_000, 2, 1),
    (2, -25, 10**9, 10**9 + 2**21, 3_600_000, 5_000_000, 6, 1),
    (2, -25, 10**9, 10**9 + 2**21, 3_600_000, 5_000_000, 1, 0),
    (10, 111, 10**12, 10**12 + 2**22, 0, 10_000_000, 6, 4),
    (10, 111, 10**12, 10**12 + 2**21, 0, 10_000_000, 1, 0),
    (2, -25, 10**6, 10**6 + 2**21, 3_600_000, 5_000_000, 2, 1),
    (3, -2, 4 * 10**18, 4 * 10**18 + 2**20, 0, 10_000_000, 1, 0),
]


@pytest.mark.parametrize("classes", [True, False])
@pytest.mark.parametrize("it", ITEMS)
def test_relations_match_oracle(gpu, mps, oracle, it, classes):
    b, c, P0, P1, N0, N1, mod, res = it
    print(f"relations {it} classes={classes}", flush=True)
    gpu.set_classes(classes)
    try:
        rels, st = gpu.run_item(mps.work_item(b, c, P0, P1, N0, N1, mod, res), chunk_ints=1 << 20)
    finally:
        gpu.set_classes(True)
    want, _ = oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, mod, res), P0, P1, 1 << 20, 8)
    assert rels == want
    assert st.declined == 0


def test_kaggle_self_test_item(gpu, mps, oracle):
    """kaggle/kaggle_run.py:66: GPU-only relations equal the CPU ones and there is at least one."""
    P0, P1 = 10_000_000_000, 10_020_000_000
    item = mps.work_item(2, -25, P0, P1, 3_600_000, 5_000_000, 2, 1)
    rels, _ = gpu.run_item(item, chunk_ints=1 << 24)
    want, _ = oracle.process_range(oracle.make_item(2, -25, P0, P1, 3_600_000, 5_000_000, 2, 1), P0, P1, 1 << 22, 8)
    assert rels == want
    assert len(rels) > 1


@pytest.mark.parametrize("field", [0, 1])
def test_fields_agree(gpu, mps, oracle, field):
    """The Montgomery and FP64 fields give the same relations below 2^50."""
    P0, P1 = 10**12, 10**12 + 2**21
    gpu.configure(field, 0)
    try:
        rels, _ = gpu.run_item(mps.work_item(10, 111, P0, P1, 0, 10_000_000, 6, 4), chunk_ints=1 << 21)
    finally:
        gpu.configure(-1, 0)
    want, _ = oracle.process_range(oracle.make_item(10, 111, P0, P1, 0, 10_000_000, 6, 4), P0, P1, 1 << 20, 8)
    assert rels == want


def test_filter_path_matches_oracle(gpu, mps, oracle):
    """GpuSiever::filter (Gpu.hpp:726-798): CPU-supplied primes, flagged indices."""
    P0, P1 = 10**9, 10**9 + 2**21
    item = mps.work_item(2, -25, P0, P1, 3_600_000, 5_000_000, 2, 1)
    primes = oracle.sieve_primes(P0, P1)
    todo = []
    assert gpu.filter(primes, item, todo)
    assert todo == sorted(todo)
    rels = []
    for i in todo:
        r = mps.run_bsgs(int(primes[i]), item)
        if r:
            rels.append(r)
    want, _ = oracle.process_range(oracle.make_item(2, -25, P0, P1, 3_600_000, 5_000_000, 2, 1), P0, P1, 1 << 20, 8)
    assert rels == want


def test_filter_full_chunk_is_exact_and_repeatable(gpu, mps, oracle):
    """GpuSiever::filter on a whole 2^26-integer chunk (3.2M primes, about half of them BSGS
    entries): the passes fit the device's entry list, so nothing overflows to the host, the flagged
    indices are the same on every call, and they carry exactly the oracle's relations."""
    P0, P1 = 10**9, 10**9 + 2**26
    item = mps.work_item(2, -25, P0, P1, 3_600_000, 5_000_000, 2, 1)
    primes = oracle.sieve_primes(P0, P1)
    first, second = [], []
    assert gpu.filter(primes, item, first), gpu.last_error()
    assert gpu.filter(primes, item, second), gpu.last_error()
    assert first == second
    assert len(first) < 4 * 2224  # the relations plus the few left to the host
    rels = [r for r in (mps.run_bsgs(int(primes[i]), item) for i in first) if r]
    want, _ = oracle.process_range(oracle.make_item(2, -25, P0, P1, 3_600_000, 5_000_000, 2, 1), P0, P1, 1 << 22, 8)
    assert len(want) == 2224
    assert rels == want


@pytest.mark.parametrize("it", ITEMS[:5])
def test_device_relations_exact(gpu, mps, oracle, it):
    """The device's own verdicts are exact: every record it decides is an oracle relation with
    the same M (smallest in-class n), and every oracle relation appears."""
    b, c, P0, P1, N0, N1, mod, res = it
    print(f"device relations {it}", flush=True)
    item = mps.work_item(b, c, P0, P1, N0, N1, mod, res)
    recs = gpu.sieve_relations(P0, P1, item)
    assert recs is not None
    want, _ = oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, mod, res), P0, P1, 1 << 20, 8)
    want_pm = {(p, M) for p, M, q in want}
    decided = {(p, M) for p, M in recs if M != mps.HOST_DECIDES}
    host = {p for p, M in recs if M == mps.HOST_DECIDES}
    assert decided <= want_pm
    assert {p for p, _ in want_pm} <= {p for p, _ in decided} | host
    assert len(host) <= max(2, len(want) // 10)


@pytest.mark.parametrize("it", ITEMS)
def test_order_mode_relations_match_oracle(gpu, mps, oracle, it):
    """Class policy 2 (one exponent class per prime from the power residues of b and -c, solved
    by Pohlig-Hellman on the device) gives the oracle's relations."""
    b, c, P0, P1, N0, N1, mod, res = it
    print(f"order mode {it}", flush=True)
    gpu.set_class_policy(2)
    try:
        rels, st = gpu.run_item(mps.work_item(b, c, P0, P1, N0, N1, mod, res), chunk_ints=1 << 20)
    finally:
        gpu.set_class_policy(1)
    want, _ = oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, mod, res), P0, P1, 1 << 20, 8)
    assert rels == want
    assert st.declined == 0


# Order-mode setup raises b and t = -c to e = (p - 1) / g with exact small-integer multiplies
# while b * p and |c| * p stay below 2^52, and with full products above; these ranges straddle
# that bound (p = 2^52 / 1000003, about 4.5e9), with relations on both sides.
SMALL_BASE_EDGE = (1 << 52) // 1000003
SMALL_BASE_ITEMS = [
    (3, -1000003, SMALL_BASE_EDGE - 2**20, SMALL_BASE_EDGE + 2**20, 0, 10_000_000, 6, 1),
    (1000003, 5, SMALL_BASE_EDGE - 2**20, SMALL_BASE_EDGE + 2**20, 0, 10_000_000, 1, 0),
]


@pytest.mark.parametrize("it", SMALL_BASE_ITEMS)
def test_order_mode_small_base_edge(gpu, mps, oracle, it):
    b, c, P0, P1, N0, N1, mod, res = it
    print(f"order mode small-base edge {it}", flush=True)
    gpu.set_class_policy(2)
    try:
        rels, st = gpu.run_item(mps.work_item(b, c, P0, P1, N0, N1, mod, res), chunk_ints=1 << 20)
    finally:
        gpu.set_class_policy(1)
    want, _ = oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, mod, res), P0, P1, 1 << 20, 8)
    assert rels == want
    assert st.declined == 0
    assert any(p < SMALL_BASE_EDGE for p, _, _ in want) and any(p > SMALL_BASE_EDGE for p, _, _ in want)


# Class moduli past the primes and prime powers of 367567200 = 2^5 3^3 5^2 7 11 13 17: 38 = 2 * 19
# and 46 = 2 * 23 bring in primes that the prefix products of Qo (x 19, 23, 29, 31 up to 2^52)
# hold as well, and 81 = 3^4 a finer 3-part than Qo's own.  Order mode runs one work-list pair per
# prime of g.
ORDER_CM_ITEMS = [
    (2, -25, 10**6, 10**6 + 2**21, 3_600_000, 5_000_000, 38, 5),
    (2, -25, 10**6, 10**6 + 2**21, 3_600_000, 5_000_000, 81, 7),
    (10, 111, 10**9, 10**9 + 2**21, 0, 10_000_000, 38, 4),
    (3, -2, 10**9, 10**9 + 2**21, 0, 10_000_000, 81, 7),
    (7, -3, 10**9, 10**9 + 2**21, 0, 10_000_000, 46, 5),
]


@pytest.mark.parametrize("it", ORDER_CM_ITEMS)
def test_order_mode_wide_class_moduli(gpu, mps, oracle, it):
    b, c, P0, P1, N0, N1, mod, res = it
    print(f"order mode, class modulus {mod}: {it}", flush=True)
    gpu.set_class_policy(2)
    try:
        rels, st = gpu.run_item(mps.work_item(b, c, P0, P1, N0, N1, mod, res), chunk_ints=1 << 20)
    finally:
        gpu.set_class_policy(1)
    want, _ = oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, mod, res), P0, P1, 1 << 20, 8)
    assert want and rels == want
    assert st.declined == 0


# base-2 Fermat pseudoprimes near 1e12 (composites that pass the device's first primality test)
PSEUDOPRIMES = [1000254486421, 1000001633401, 1000133402761]


@pytest.mark.parametrize("policy", [1, 2])
@pytest.mark.parametrize("psp", PSEUDOPRIMES)
def test_pseudoprime_candidates(gpu, mps, oracle, policy, psp):
    """A base-2 pseudoprime reaches the setup stage; it must be dropped or left to the host, never
    stall the kernel, and the relations around it must be the oracle's."""
    assert pow(2, psp - 1, psp) == 1
    P0, P1 = psp - 2**19, psp + 2**19
    item = (10, 111, P0, P1, 0, 10_000_000, 6, 4)
    print(f"pseudoprime {psp} policy {policy}", flush=True)
    gpu.set_class_policy(policy)
    try:
        rels, _ = gpu.run_item(mps.work_item(*item), chunk_ints=1 << 21)
    finally:
        gpu.set_class_policy(1)
    want, _ = oracle.process_range(oracle.make_item(*item), P0, P1, 1 << 20, 8)
    assert rels == want


@pytest.mark.parametrize("policy", [1, 2])
@pytest.mark.parametrize("psp", [1000254486421, 1000001376901])
def test_pseudoprime_candidates_base2_item(gpu, mps, oracle, policy, psp):
    """With b = 2 the order-mode Fermat test (to base b) is the base-2 test, so these pseudoprimes
    (p = 5 mod 8: (2/p) = -1, kept by the filter when cm = 1) pass it and reach the discrete logs,
    which must drop them; the relations around them must be the oracle's."""
    assert pow(2, psp - 1, psp) == 1 and psp % 8 == 5
    P0, P1 = psp - 2**19, psp + 2**19
    item = (2, 3, P0, P1, 0, 10_000_000, 1, 0)
    print(f"pseudoprime {psp} b=2 policy {policy}", flush=True)
    gpu.set_class_policy(policy)
    try:
        rels, _ = gpu.run_item(mps.work_item(*item), chunk_ints=1 << 21)
    finally:
        gpu.set_class_policy(1)
    want, _ = oracle.process_range(oracle.make_item(*item), P0, P1, 1 << 20, 8)
    assert rels == want


@pytest.mark.parametrize("policy", [1, 2])
def test_pseudoprime_with_a_solution_is_dropped(gpu, mps, oracle, policy):
    """A base-2 pseudoprime q that divides b^n + c for an in-class n (c = -(2^1000003 mod q)) gets
    through the Fermat tests and has a BSGS solution: only the Miller-Rabin test over the records
    removes it.  The relations must be the oracle's, which has no composite."""
    q = 1000254486421
    assert pow(2, q - 1, q) == 1 and q % 8 == 5
    c = -pow(2, 1000003, q)
    P0, P1 = q - 2**19, q + 2**19
    item = (2, c, P0, P1, 1_000_000, 1_100_000, 1, 0)
    print(f"pseudoprime with solution, policy {policy}", flush=True)
    gpu.set_class_policy(policy)
    try:
        rels, _ = gpu.run_item(mps.work_item(*item), chunk_ints=1 << 21)
    finally:
        gpu.set_class_policy(1)
    want, _ = oracle.process_range(oracle.make_item(*item), P0, P1, 1 << 20, 8)
    assert q not in [r[0] for r in rels]
    assert rels == want


# primes of [1001, 1001 + 2^20] where 2 has order below 32 (divisors of 2^k - 1, k < 32)
SMALL_ORDER = {1103: 29, 1801: 25, 2089: 29, 2731: 26, 8191: 13, 131071: 17, 178481: 23, 262657: 27, 524287: 19}


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_small_order_primes_left_to_host(gpu, mps, oracle, policy):
    """b = 2, c = -1: p | 2^n - 1 exactly when ord_p(2) divides n, so every prime has a relation.
    Where the BSGS base has order below the baby-step count (the primes above, with at least 64
    baby steps per entry) the baby values repeat and the smallest solution might be missed; the
    device must leave those primes to the host, and the relations must be the oracle's."""
    P0, P1 = 1001, 1001 + 2**20
    item = (2, -1, P0, P1, 1, 100_000, 1, 0)
    print(f"small-order primes, policy {policy}", flush=True)
    gpu.set_class_policy(policy)
    try:
        recs = gpu.sieve_relations(P0, P1, mps.work_item(*item))
        rels, _ = gpu.run_item(mps.work_item(*item), chunk_ints=1 << 21)
    finally:
        gpu.set_class_policy(1)
    assert recs is not None
    host = {p for p, M in recs if M == mps.HOST_DECIDES}
    assert set(SMALL_ORDER) <= host
    want, _ = oracle.process_range(oracle.make_item(*item), P0, P1, 1 << 20, 8)
    assert rels == want
    got = {p: M for p, M, q in rels}
    assert all(got[p] == o for p, o in SMALL_ORDER.items())


# Class moduli whose Qo order mode cannot take: a prime factor of 2^16 or more (2 * 65537) or a
# prime power of 2^16 or more (2^16, 3^11).  Under class policy 2 such an item runs in class or
# plain mode like any other; it must not be declined.
UNSUPPORTED_CM = [(2 * 65537, 3), (65536, 5), (3**11, 7)]


@pytest.mark.parametrize("mod,res", UNSUPPORTED_CM)
def test_order_mode_unsupported_modulus_falls_back(gpu, mps, oracle, mod, res):
    P0, P1 = 10**6, 10**6 + 2**21
    item = (2, -25, P0, P1, 0, 10_000_000, mod, res)
    print(f"policy 2, class modulus {mod}", flush=True)
    gpu.set_class_policy(2)
    try:
        rels, st = gpu.run_item(mps.work_item(*item), chunk_ints=1 << 20)
    finally:
        gpu.set_class_policy(1)
    want, _ = oracle.process_range(oracle.make_item(*item), P0, P1, 1 << 20, 16)
    assert st.declined == 0
    assert want and rels == want


INT64_MIN, INT64_MAX = -(1 << 63), (1 << 63) - 1
EXTREME_C = [
    (2, INT64_MIN, 10**4, 10**4 + 2**16, 0, 10**6, 1, 0),
    (3, INT64_MAX, 10**5, 10**5 + 2**16, 0, 10**6, 2, 1),
    (5, INT64_MIN, 3 * 10**5, 3 * 10**5 + 2**16, 10**5, 2 * 10**6, 6, 5),
]


@pytest.mark.parametrize("mode", [(False, 1, -1), (True, 1, -1), (True, 2, -1), (True, 2, 0)],
                         ids=["plain", "class", "order", "order-mont"])
@pytest.mark.parametrize("it", EXTREME_C)
def test_extreme_constants(gpu, mps, oracle, it, mode):
    """c at the ends of int64: t = -c is 2^63 for c = INT64_MIN, which a 64-bit negation would turn
    into -2^63 (the wrong residue, and the wrong quadratic character for p == 3 mod 4)."""
    classes, policy, field = mode
    print(f"extreme c {it} mode {mode}", flush=True)
    gpu.set_classes(classes)
    gpu.set_class_policy(policy)
    gpu.configure(field, 0)
    try:
        rels, st = gpu.run_item(mps.work_item(*it), chunk_ints=1 << 16)
    finally:
        gpu.set_classes(True)
        gpu.set_class_policy(1)
        gpu.configure(-1, 0)
    assert st.declined == 0
    want, _ = oracle.process_range(oracle.make_item(*it), it[2], it[3], 1 << 16, 8)
    assert want
    assert rels == want


EDGE_ITEMS = [
    (0, -5, 10**4, 10**4 + 2**14, 0, 1000, 1, 0),  # b = 0: every prime divides b, the host decides
    (1, -1, 10**4, 10**4 + 2**14, 10, 1000, 1, 0),  # b = 1, c = -1: every prime, M = N0, q = 1
    (1, 4, 10**4, 10**4 + 2**14, 0, 1000, 1, 0),  # b = 1: only primes dividing 1 + c
    (-3, 7, 10**5, 10**5 + 2**16, 0, 10**5, 2, 1),  # negative base
    (2, -25, 10**5, 10**5 + 2**16, -100, 10**5, 7, 12),  # N0 < 0, cls_res >= cls_mod
    (3, -2, 10**5, 10**5 + 2**16, 0, 10**6, 1 << 40, 5),  # one in-class exponent
    (12, 5, 10**12, 10**12 + 2**16, 0, 10**7, 131074, 3),  # cm = 2 * 65537
    (10, -1, 10**6, 10**6 + 2**16, 5, 5, 1, 0),  # empty exponent range
]


@pytest.mark.parametrize("mode", [(False, 1, -1), (True, 1, -1), (True, 2, -1), (True, 2, 0)],
                         ids=["plain", "class", "order", "order-mont"])
@pytest.mark.parametrize("it", EDGE_ITEMS)
def test_edge_items(gpu, mps, oracle, it, mode):
    """Degenerate and boundary work items, in every setup mode, against the oracle; none of them
    may be declined."""
    classes, policy, field = mode
    print(f"edge item {it} mode {mode}", flush=True)
    gpu.set_classes(classes)
    gpu.set_class_policy(policy)
    gpu.configure(field, 0)
    try:
        rels, st = gpu.run_item(mps.work_item(*it), chunk_ints=1 << 16)
    finally:
        gpu.set_classes(True)
        gpu.set_class_policy(1)
        gpu.configure(-1, 0)
    assert st.declined == 0
    want, _ = oracle.process_range(oracle.make_item(*it), it[2], it[3], 1 << 16, 8)
    assert rels == want


MULTIPASS_ITEMS = [
    (2, -3, 5, 2**16, 0, 1000, 1, 0),  # from the smallest allowed lo: primes below 1000 go to the host
    (2, -25, 10**9, 10**9 + 2**20, 3_600_000, 5_000_000, 2, 1),
    (2, -25, 10**9 + 7, 10**9 + 8, 0, 10_000_000, 1, 0),  # one odd number
    (2, -25, 10**9 + 8, 10**9 + 9, 0, 10_000_000, 1, 0),  # no odd number
]


@pytest.mark.parametrize("pass_odds", [4096, 5000, (1 << 17) + 12345])
def test_chunks_of_many_passes(gpu, mps, oracle, pass_odds):
    """Chunks cut into many device passes (pass sizes that are not multiples of the 2^17-odd sieve
    segment), windows of one odd number or none, the smallest lo and the top of the 62-bit range
    (a planted relation at p ~ 4.6e18): the relations equal the oracle's."""
    import json
    import os
    planted = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "planted.json")))["items"]
    items = MULTIPASS_ITEMS + [tuple(it["item"]) for it in planted if it["tag"].startswith("top62")]
    print(f"pass_odds {pass_odds}", flush=True)
    assert gpu.configure(-1, pass_odds) == mps.MPS_OK
    try:
        for it in items:
            rels, st = gpu.run_item(mps.work_item(*it), chunk_ints=1 << 18)
            assert st.declined == 0, it
            want, _ = oracle.process_range(oracle.make_item(*it), it[2], it[3], 1 << 18, 8)
            assert rels == want, it
    finally:
        gpu.configure(-1, 1 << 25)


SPLIT_ITEMS = [
    (2, -25, 10**9, 10**9 + 2**22, 3_600_000, 5_000_000, 2, 1),  # relations at small p
    (10, 111, 10**12, 10**12 + 3 * 2**20 + 5, 0, 10**7, 6, 4),  # c2's item, a short odd-sized window
]


@pytest.mark.parametrize("it", SPLIT_ITEMS, ids=["c1-like", "c2-like"])
def test_small_item_split_among_feeders(gpu, mps, oracle, it):
    """An item of fewer 2^26-integer chunks than feeder threads is split evenly among the feeders
    (each on its own pass lane): the relations equal those of one feeder running the single chunk,
    and the oracle's."""
    print(f"split item {it}", flush=True)
    w = mps.work_item(*it)
    rels2, st2 = gpu.run_item(w)  # two feeders: two chunks
    rels1, st1 = gpu.run_item(w, feeder_threads=1)  # one chunk
    assert st2.declined == 0 and st1.declined == 0
    assert st2.chunks == 2 and st1.chunks == 1
    want, _ = oracle.process_range(oracle.make_item(*it), it[2], it[3], 1 << 18, 8)
    assert rels2 == rels1 == want
    assert it[0] != 2 or len(want) > 0
