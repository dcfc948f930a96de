"""Stage-level parity of the HIP kernels against the oracle's restatement of each device stage.

Sieve survivors must equal the oracle's psieve set (Gpu.hpp:385-455); BSGS entries must equal the
prime set (and class masks) the oracle's prep_core / prep_q_core decide (Gpu.hpp:231-308,
:478-589).  Both fields (FP64 below 2^50, Montgomery everywhere) are covered.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MCAP = 2048  # default largest baby-step count without classes

SIEVE_RANGES = [
    (5, 5 + 2 * 150_000),                  # from the smallest start the device takes: sieving primes survive
    (2**15 + 1, 65536 + 2**18),            # across the largest sieving prime
    (1001, 1001 + 2 * 300_000),            # tiny p: survivors are the primes
    (10**9, 10**9 + 2**22),                # c1 scale: sqrt(hi) < 65521, survivors are the primes
    (10**12 + 12345, 10**12 + 12345 + 2**23),
    (95 * 10**12, 95 * 10**12 + 2**21),
    (4 * 10**18, 4 * 10**18 + 2**22),      # near the 62-bit limit
    ((1 << 62) - 2**21, (1 << 62) - 1),
]


@pytest.mark.parametrize("lo,hi", SIEVE_RANGES)
def test_sieve_survivors_match_oracle(gpu, mps, oracle, lo, hi):
    print(f"sieve [{lo}, {hi})", flush=True)
    item = mps.work_item(2, -25, lo, hi, 0, 1000, 2, 1)
    got = np.sort(gpu.debug_stage(0, lo, hi, item))
    want = oracle.sieve_survivors(lo, hi)
    assert len(got) == len(want)
    assert np.array_equal(got, want)


ITEMS = [
    # (b, c, lo, hi, N0, N1, mod, res, classes)
    (2, -25, 10**9, 10**9 + 2**20, 3_600_000, 5_000_000, 2, 1, False),
    (2, -25, 10**9, 10**9 + 2**20, 3_600_000, 5_000_000, 2, 1, True),
    (10, 111, 10**12, 10**12 + 2**20, 0, 10_000_000, 6, 4, True),
    (10, 111, 10**12, 10**12 + 2**20, 0, 10_000_000, 1, 0, True),
    (2, -25, 95 * 10**12, 95 * 10**12 + 2**20, 3_600_000, 5_000_000, 6, 1, True),
    (3, -2, 4 * 10**18, 4 * 10**18 + 2**19, 0, 10_000_000, 1, 0, True),
    (7, 3, 10**6 + 1, 10**6 + 2**19, 12345, 412345, 2, 0, True),
]


def class_modulus(oracle, oit, mod, N0, N1, classes, policy):
    """(class mode?, Q) of the device: the reference's gpu_classes (Math.hpp:324-337), which
    needs an even class modulus, or under class policy 1 the same rule for any modulus."""
    qmode, Q, _ = oracle.gpu_classes(oit)
    if not qmode and policy == 1 and N1 > N0 >= 0:
        cm = max(1, mod)
        Q = cm // np.gcd(cm, 24) * 24
        qmode = Q <= 48 and (N1 - N0) // Q >= 4096
    return qmode and classes, int(Q)


def offset_order(mask, qp, mod, res, N0):
    """The oracle's class mask (bit j: residue cres + j*cm modulo qp, prep_q_core's numbering,
    Gpu.hpp:491-498) in the device's numbering: bit k for the class whose first exponent lies k*cm
    past n_start, the first in-class exponent >= N0."""
    cm = max(1, mod)
    N0 = max(0, N0)
    off0 = (res % cm - N0) % cm  # n_start - N0
    out = 0
    for j in range(qp // cm):
        if (mask >> j) & 1:
            dj = (res % cm + j * cm - N0) % qp  # offset of class j's first exponent from N0
            assert (dj - off0) % cm == 0
            out |= 1 << ((dj - off0) // cm)
    return out


@pytest.mark.parametrize("field,policy", [(-1, 1), (0, 1), (-1, 0)])
@pytest.mark.parametrize("it", ITEMS)
def test_entries_match_oracle_decisions(gpu, mps, oracle, it, field, policy):
    b, c, lo, hi, N0, N1, mod, res, classes = it
    if field == 0 or hi > (1 << 50):
        field_name = "montgomery"
    else:
        field_name = "fp64"
    print(f"entries {it} field={field_name} class policy {policy}", flush=True)
    gpu.configure(field, 0)
    gpu.set_classes(classes)
    gpu.set_class_policy(policy)
    try:
        item = mps.work_item(b, c, lo, hi, N0, N1, mod, res)
        ent = gpu.debug_stage(2, lo, hi, item)
    finally:
        gpu.configure(-1, 0)
        gpu.set_classes(True)
        gpu.set_class_policy(1)
    oit = oracle.make_item(b, c, lo, hi, N0, N1, mod, res)
    qmode, Q = class_modulus(oracle, oit, mod, N0, N1, classes, policy)
    surv = oracle.sieve_survivors(lo, hi)
    want = {}
    for x in surv.tolist():
        if qmode:
            r, e = oracle.prep_q(x, oit, Q, True)
            if r == 1:
                want[x] = (e.s, offset_order(e.mask, e.qp, mod, res, N0))
        else:
            r, e = oracle.prep(x, oit, True)
            if r == 1:
                want[x] = (e.jcount, e.stride, e.parity)
    got = {}
    for p, m, ng, sinfo in ent.tolist():
        s, info = sinfo & 0xFF, sinfo >> 8
        if qmode:
            got[p] = (s, info)
        else:
            # the device picks its own split (m <= MCAP baby steps, ng giant steps g = 0 .. ng-1):
            # every j in [0, jcount) must be g*m - i for some g < ng, i < m
            jc = want.get(p, (0,))[0]
            assert 1 <= m <= MCAP and m * (ng - 1) >= jc, (p, m, ng, jc)
            got[p] = (jc, 2 if info & 1 else 1, (info >> 1) & 1)
    assert len(got) == len(ent), "an entry was written twice"
    assert set(got) == set(want)
    assert got == want


ORDER_Q = 367567200  # 2^5 3^3 5^2 7 11 13 17
ORDER_MORE = (19, 23, 29, 31, 37)


def order_modulus(cm):
    """Qo = lcm(cm, B) for the longest prefix product B of 367567200, then times 19, 23, 29, 31,
    37, with Qo below 2^52."""
    import math
    lcm = lambda a, b: a // math.gcd(a, b) * b
    B = ORDER_Q
    Qo = lcm(cm, B)
    for r in ORDER_MORE:
        B *= r
        if lcm(cm, B) >= 2**52:
            break
        Qo = lcm(cm, B)
    return Qo


# the library's defaults for the medium primes of p - 1 (order mode): the bound of the Montgomery
# field's passes (the FP64 field's use none unless a tuning sets a bound) and the smallest class
# they narrow
EXT_BOUND, EXT_KMIN = 1024, 131072


def ext_primes(Qo, bound):
    """The medium primes the device tries: the primes from 19 on, from the first that does not
    divide Qo, up to bound, at most 192 of them."""
    import sympy
    ps = list(sympy.primerange(19, 2000))
    i = 0
    while i < len(ps) and Qo % ps[i] == 0:
        i += 1
    return [r for r in ps[i:i + 192] if r <= bound]


def order_class(p, b, c, cm, cres, N0, N1, ext_bound=EXT_BOUND):
    """The exponent class the device must find in order mode, restated from the mathematics: with
    g = gcd(Qo, p - 1) (order_modulus), every solution n of b^n == -c (mod p) has (b^e)^n == (-c)^e,
    e = (p-1)/g; per prime power of g this fixes n modulo the order of the projected b^e (or rules
    p out), and with n == cres (mod cm) the admissible n are one class mod L; a class of at least
    EXT_KMIN exponents is then narrowed by the medium primes of p - 1 (below).  Returns
    (L, first n >= N0, count below N1) or None."""
    import math
    import sympy
    Qo = order_modulus(cm)
    g = math.gcd(Qo, p - 1)
    e = (p - 1) // g
    Be, Te = pow(b % p, e, p), pow((-c) % p, e, p)
    residues = []  # (k_r, d_r)
    for r in sympy.primefactors(g):
        ra = r ** sympy.multiplicity(r, g)
        x, y = pow(Be, g // ra, p), pow(Te, g // ra, p)
        d = next(k for k in range(1, ra + 1) if pow(x, k, p) == 1)
        ks = [k for k in range(d) if pow(x, k, p) == y]
        if not ks:
            return None
        residues.append((ks[0], d))
    residues.append((cres % cm, cm))
    # the CRT over moduli that share primes (sympy's general solver; no code shared with the
    # device's CRT basis of Qo)
    from sympy.ntheory.modular import solve_congruence
    sol = solve_congruence(*residues)
    if sol is None:
        return None
    n0, L = int(sol[0]), int(sol[1])
    first = N0 + (n0 - N0) % L
    K = (N1 - 1 - first) // L + 1 if first < N1 else 0
    if K <= 0:
        return None
    if ext_bound is None or min(K, p) < EXT_KMIN:
        return (L, first, K)
    # medium primes r of p - 1 (ascending, at most 8, while L times their product stays below
    # 2^62): every solution has x^n == y for x = b^((p-1)/r), y = (-c)^((p-1)/r); x of order r fixes
    # n mod r, x == 1 rules p out unless y == 1
    taken, lp = [], L
    for r in ext_primes(Qo, ext_bound):
        if (p - 1) % r == 0 and lp <= ((2**64 - 1) // r) >> 2 and len(taken) < 8:
            taken.append(r)
            lp *= r
    for r in taken:
        x, y = pow(b % p, (p - 1) // r, p), pow((-c) % p, (p - 1) // r, p)
        if x == 1:
            if y != 1:
                return None
            continue
        if L % r == 0:
            continue
        k, v = 0, 1
        while v != y and k < r:
            v, k = v * x % p, k + 1
        assert k < r  # x has prime order r and y is a power of x (p is prime)
        n0, L = (int(v) for v in solve_congruence((first, L), (k, r)))
        first = N0 + (n0 - N0) % L
        if first >= N1:
            return None
    return (L, first, (N1 - 1 - first) // L + 1)


ORDER_ITEMS = [
    (10, 111, 10**12, 10**12 + 2**19, 0, 10_000_000, 6, 4),
    (2, -25, 10**9, 10**9 + 2**19, 3_600_000, 5_000_000, 2, 1),
    (3, -2, 4 * 10**18, 4 * 10**18 + 2**18, 0, 10_000_000, 1, 0),
    (7, 3, 10**6 + 1, 10**6 + 2**19, 12345, 412345, 1, 0),
]


@pytest.mark.parametrize("ext", [0, 1024, 1, 300])
@pytest.mark.parametrize("it", ORDER_ITEMS)
def test_order_mode_entries_match_power_residues(gpu, mps, oracle, it, ext):
    """Under class policy 2 the prime entries are exactly the primes whose power-residue class is
    non-empty (narrowed by the medium primes of p - 1 where they are on: by default in the
    Montgomery field's passes, p > 2^50; in both fields with an explicit bound), and each entry's
    baby/giant split covers the class."""
    import sympy
    b, c, lo, hi, N0, N1, mod, res = it
    print(f"order entries {it} ext {ext}", flush=True)
    w = mps.work_item(b, c, lo, hi, N0, N1, mod, res)
    gpu.set_tuning(w, policy=2, ext=ext)
    try:
        ent = gpu.debug_stage(2, lo, hi, w)
    finally:
        gpu.clear_tuning()
    bound = (EXT_BOUND if hi > 2**50 else None) if ext == 0 else (None if ext == 1 else ext)
    nums = [int(p) for p, _, _, _ in ent.tolist()]
    assert len(set(nums)) == len(nums), "an entry was written twice (or a slot kept an old entry)"
    got = {int(p): (int(m), int(ng)) for p, m, ng, _ in ent.tolist() if sympy.isprime(int(p))}
    want = {}
    for x in oracle.sieve_primes(lo, hi).tolist():
        if x < 1000 or b % x == 0 or c % x == 0:
            continue
        oc = order_class(x, b, c, max(1, mod), res, N0, N1, bound)
        if oc:
            want[x] = min(oc[2], x)
    assert set(got) == set(want)
    for p, (m, ng) in got.items():
        assert m * (ng - 1) >= want[p], p
