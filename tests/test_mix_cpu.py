"""The seeded serving mix (mass-prime-siever_amd/mix.py): the committed fixture is the stream the
generator draws for seed 0, every item lies in the stated ranges, no item is identically
reducible (checked here by direct computation, not by the generator's helpers), and the prime
counts of a few items agree with sympy."""
import json
import math
import os
import sys

import pytest
import sympy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "mass-prime-siever_amd"))
import mix  # noqa: E402

FIXTURE = json.load(open(os.path.join(HERE, "golden", "mix_seed0.json")))


def test_fixture_is_seed0_stream():
    items = mix.serving_mix(0, 256)
    fx = FIXTURE["items"]
    assert len(fx) == 256
    for a, b in zip(items, fx):
        assert {k: b[k] for k in a} == a


def test_ranges():
    for it in FIXTURE["items"]:
        assert it["b"] in (2, 3, 5, 6, 7, 10, 12)
        assert it["c"] != 0 and -9999 <= it["c"] <= 9999
        assert 10**9 <= it["P0"] < 10**15
        assert 2**24 <= it["P1"] - it["P0"] < 2**30
        assert 0 <= it["N0"] < 5_000_000
        assert 10**5 <= it["N1"] - it["N0"] < 10**7
        assert it["mod"] in (1, 2, 6) and 0 <= it["res"] < max(1, it["mod"])
        assert it["gap"] > 0


def test_not_identically_reducible():
    for it in FIXTURE["items"]:
        b, c, mod, res = it["b"], it["c"], it["mod"], it["res"]
        assert math.gcd(b, c) == 1
        # no prime below 100 divides b^n + c for every n (a fixed divisor divides b - 1 and 1 + c)
        for r in sympy.primerange(2, 100):
            assert any((pow(b, n, r) + c) % r for n in range(r)), (b, c, r)
        ns = [n for n in range(it["N0"], it["N0"] + 60) if mod == 1 or n % mod == res]
        if all(n % 2 == 0 for n in ns):
            assert not (c < 0 and sympy.perfect_power(-c) and math.isqrt(-c) ** 2 == -c)


@pytest.mark.parametrize("k", [0, 7, 100])
def test_prime_counts(k):
    """The counting sieve behind the fixture's counts agrees with sympy.isprime on a slice of the
    item's range (the whole range is too long for sympy here)."""
    it = FIXTURE["items"][k]
    lo, hi = it["P0"], it["P0"] + 2**14
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
    import oracle as O
    assert O.lib().or_sieve_primes(lo, hi, None, 0) == sum(1 for x in range(lo | 1, hi, 2) if sympy.isprime(x))
    assert it["primes"] > 0
