"""Pins of the CPU oracle (oracle/mps_oracle.c) against checks that share no code with it:
closed forms in Python integers, sympy, brute force over every exponent, and the brute-force
golden fixtures of tests/golden/relations_small.json."""
import json
import math
import os
import random

import pytest
import sympy
from hypothesis import given, settings, strategies as st

HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(HERE, "golden", "relations_small.json")) as f:
    GOLDEN = json.load(f)


def test_mont_mul_closed_form(oracle):
    """Montgomery product (Math.hpp:47-61) == a*b*2^-64 mod n, for n up to 2^62."""
    rng = random.Random(1)
    for bits in (11, 30, 40, 50, 61, 62):
        for _ in range(200):
            n = rng.randrange(2 ** (bits - 1), 2 ** bits) | 1
            a, b = rng.randrange(n), rng.randrange(n)
            assert oracle.mont_mul(a, b, n) == a * b * pow(2 ** 64, -1, n) % n


def test_is_prime64_known_answers(oracle):
    """is_prime64 (Math.hpp:209-251): exact for every 64-bit n with the 12 bases 2..37."""
    # strong pseudoprimes: 2047 to base 2; 3825123056546413051 to bases 2..23 (needs 29, 31 or 37)
    for n in (0, 1, 4, 561, 1105, 2047, 1373653, 25326001, 3215031751, 2152302898747, 3474749660383,
              341550071728321, 3825123056546413051, 318665857834031151167461 % (1 << 64)):
        assert oracle.is_prime64(n) == sympy.isprime(n), n
    for n in (2, 3, 37, 1000000007, 2 ** 61 - 1, (1 << 62) - 57, 18446744073709551557):
        assert oracle.is_prime64(n)


def test_is_prime64_random(oracle):
    rng = random.Random(2)
    for _ in range(3000):
        n = rng.randrange(2 ** rng.randrange(2, 64))
        assert oracle.is_prime64(n) == sympy.isprime(n), n


def test_multiplicative_order(oracle):
    """multiplicative_order (Math.hpp:294-310) == sympy.n_order, including p - 1 with large factors."""
    rng = random.Random(3)
    for bits in (12, 30, 45, 61):
        for _ in range(40):
            p = sympy.randprime(2 ** (bits - 1), 2 ** bits)
            b = rng.randrange(2, min(p, 10 ** 6))
            assert oracle.multiplicative_order(b, p) == sympy.n_order(b, p)
    p = 2 * 2305843009213693951 + 1  # p - 1 = 2 * (2^61 - 1) when p is prime
    if sympy.isprime(p):
        assert oracle.multiplicative_order(3, p) == sympy.n_order(3, p)


def test_jacobi(oracle):
    rng = random.Random(4)
    for _ in range(2000):
        n = rng.randrange(3, 2 ** 62) | 1
        a = rng.randrange(0, 2 ** 62)
        assert oracle.jacobi(a, n) == sympy.jacobi_symbol(a, n)


def test_sieve_primes(oracle):
    """segmented_sieve (Sieve.hpp:110-185): the primes of [lo, hi)."""
    for lo, hi in ((2, 1000), (5, 200000), (10 ** 9, 10 ** 9 + 300000), (2 ** 40 - 100000, 2 ** 40 + 100000)):
        got = oracle.sieve_primes(lo, hi).tolist()
        assert got == list(sympy.primerange(lo, hi))


def test_sieve_survivors_trial_division(oracle):
    """psieve's output set (Gpu.hpp:385-455): odd x with no odd prime factor q <= 65521 unless x = q."""
    qs = list(sympy.primerange(3, 65536))
    for lo, hi in ((1001, 1001 + 20000), (10 ** 12 + 1, 10 ** 12 + 20001), (4 * 10 ** 18 + 1, 4 * 10 ** 18 + 8001)):
        want = [x for x in range(lo | 1, hi, 2) if all(x % q or x == q for q in qs if q * q <= x)]
        assert oracle.sieve_survivors(lo, hi).tolist() == want


def brute_relations(b, c, P0, P1, N0, N1, mod, res):
    out = []
    for p in sympy.primerange(max(P0, 1000), P1):
        t = (-c) % p
        if b % p == 0 or t == 0:
            continue
        for n in range(max(N0, 0), N1):
            if (mod <= 1 or n % mod == res) and pow(b, n, p) == t:
                out.append((p, n, int(sympy.n_order(b, p))))
                break
    return out


@settings(max_examples=40, deadline=None)
@given(b=st.sampled_from([2, 3, 5, 6, 7, 10, 12]), c=st.integers(-200, 200).filter(lambda x: x != 0),
       P0=st.integers(1000, 30000), width=st.integers(100, 3000), N0=st.integers(0, 500), span=st.integers(1, 800),
       mod=st.sampled_from([1, 2, 3, 4, 6, 12]), res=st.integers(0, 11))
def test_run_bsgs_brute_force(oracle, b, c, P0, width, N0, span, mod, res):
    """run_bsgs (Math.hpp:377-486): M is the smallest in-class n >= N0 with p | b^n + c."""
    res %= mod
    item = oracle.make_item(b, c, P0, P0 + width, N0, N0 + span, mod, res)
    got, _ = oracle.process_range(item, P0, P0 + width, 1 << 20, 1)
    assert got == brute_relations(b, c, P0, P0 + width, N0, N0 + span, mod, res)


@pytest.mark.parametrize("rec", GOLDEN, ids=lambda r: ",".join(map(str, r["item"])))
def test_oracle_matches_golden(oracle, rec):
    b, c, P0, P1, N0, N1, mod, res = rec["item"]
    item = oracle.make_item(b, c, P0, P1, N0, N1, mod, res)
    got, _ = oracle.process_range(item, P0, P1, 1 << 20, 4)
    assert got == [tuple(r) for r in rec["relations"]]
    if "results_json" in rec:
        assert oracle.results_json(item, got, "2026-01-01 00:00:00 UTC+0") == rec["results_json"]


def residue_mask(p, b, c, N0, N1, mod, res, Qmax):
    """Classes of prep_q_core's split that can hold a solution, from the g-th power residue
    criterion: T_j = t * b^-n_j is a g-th power modulo p iff T_j^((p-1)/g) == 1."""
    g = math.gcd(Qmax, p - 1)
    qp = g // math.gcd(g, mod) * mod
    mask = 0
    for j in range(qp // mod):
        dj = (res % mod + j * mod - N0) % qp
        if dj >= N1 - N0:
            continue
        T = (-c) * pow(b, -(N0 + dj), p) % p
        if pow(T, (p - 1) // g, p) == 1:
            mask |= 1 << j
    return qp, mask


@pytest.mark.parametrize("it", [(10, 111, 10 ** 12, 0, 10 ** 7, 6, 4), (2, -25, 95 * 10 ** 12, 3600000, 5000000, 2, 1),
                                (7, 3, 10 ** 9, 12345, 412345, 2, 0), (5, -4, 10 ** 10, 1, 10 ** 6, 4, 3)])
def test_prep_q_masks_are_power_residue_classes(oracle, it):
    """prep_q_core's class masks (Gpu.hpp:478-546) for primes equal the g-th power residue classes."""
    b, c, P0, N0, N1, mod, res = it
    item = oracle.make_item(b, c, P0, P0 + 200000, N0, N1, mod, res)
    ok, Q, _ = oracle.gpu_classes(item)
    assert ok
    for p in sympy.primerange(P0, P0 + 200000):
        r, e = oracle.prep_q(p, item, Q, True)
        qp, mask = residue_mask(p, b, c, N0, N1, mod, res, Q)
        if mask == 0:
            assert r == 0, p
        else:
            assert r == 1 and e.qp == qp and e.mask == mask, p
            K = -(-(N1 - N0) // qp)
            assert (e.Gc - 1) * e.m >= K - 1


@pytest.mark.parametrize("it", [(2, -25, 10 ** 9, 3600000, 5000000, 2, 1), (3, -2, 4 * 10 ** 18, 0, 10 ** 7, 1, 0),
                                (10, 111, 10 ** 12, 0, 10 ** 7, 1, 0)])
def test_prep_plain_is_residue_filter(oracle, it):
    """prep_core (Gpu.hpp:231-308) keeps a prime iff T = t*b^-n_start lies in a coset of <B>
    that can hold a solution: for g = gcd(cm, p-1) >= 2, T is a g-th power; for g = 1, T is a
    square when b is a square, else the parity of k follows T's Legendre symbol."""
    b, c, P0, N0, N1, mod, res = it
    item = oracle.make_item(b, c, P0, P0 + 100000, N0, N1, mod, res)
    n_start = N0 + ((res % mod) + mod - (N0 % mod)) % mod
    for p in sympy.primerange(P0, P0 + 100000):
        r, e = oracle.prep(p, item, True)
        T = (-c) * pow(b, -n_start, p) % p
        g = math.gcd(mod, p - 1)
        if g >= 2:
            assert (r == 1) == (pow(T, (p - 1) // g, p) == 1), p
        else:
            qr_b = pow(b, (p - 1) // 2, p) == 1
            qr_T = pow(T, (p - 1) // 2, p) == 1
            if qr_b:
                assert (r == 1) == qr_T, p
            else:
                assert r == 1 and e.stride == 2 and e.parity == (0 if qr_T else 1), p


def test_oracle_extreme_constants(oracle):
    """c = INT64_MIN and INT64_MAX: the oracle's (p, M) equal brute force with exact integers
    (t = -c = 2^63 for INT64_MIN), and q = ord_p(b) by sympy."""
    import sympy
    for b, c, P0, P1, N0, N1 in [(2, -(1 << 63), 10**4, 10**4 + 2**11, 0, 3000),
                                 (3, (1 << 63) - 1, 10**4, 10**4 + 2**11, 0, 20000)]:
        rels, _ = oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, 1, 0), P0, P1, 1 << 12, 1)
        want = []
        for p in sympy.primerange(P0, P1):
            t, x = (-c) % p, pow(b, N0, p)
            for n in range(N0, N1):
                if x == t:
                    want.append((p, n, int(sympy.n_order(b, p))))
                    break
                x = x * b % p
        assert want and rels == want


def test_oracle_degenerate_items(oracle):
    """b = 0, b = 1 and a negative base, an N0 below 0 and cls_res >= cls_mod: the oracle's relations
    equal brute force with exact integers (M the smallest n >= max(N0, 0), n == res (mod mod),
    n < N1 with p | b^n + c; q = ord_p(b))."""
    import sympy
    items = [(0, -5, 10**4, 10**4 + 2**10, 0, 200, 1, 0), (1, -1, 10**4, 10**4 + 2**10, 10, 200, 1, 0),
             (-3, 7, 10**3 + 1, 10**3 + 2**11, 0, 3000, 2, 1), (2, -25, 10**3 + 1, 10**3 + 2**11, -100, 3000, 7, 12)]
    for b, c, P0, P1, N0, N1, mod, res in items:
        rels, _ = oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, mod, res), P0, P1, 1 << 12, 1)
        want = []
        for p in sympy.primerange(P0, P1):
            n = next((n for n in range(max(N0, 0), N1) if n % mod == res % mod and (pow(b, n, p) + c) % p == 0), None)
            if n is not None:
                want.append((p, n, int(sympy.n_order(b % p, p)) if b % p else 1))
        assert rels == want, (b, c)
