"""The planted-relation fixture (tests/golden/planted.json) on the CPU: its relations hold by
direct arithmetic, and the oracle's run_bsgs (Math.hpp:377-486) and is_prime64 (Math.hpp:209-251)
give them at every magnitude, up to p just below 2^62.  The GPU tests compare the device with the
oracle over each item's whole P window (tests/test_gpu_planted.py)."""
import json
import math
import os

import pytest
import sympy

HERE = os.path.dirname(os.path.abspath(__file__))
PLANTED = json.load(open(os.path.join(HERE, "golden", "planted.json")))["items"]


def _order_ok(b, p, q):
    if pow(b, q, p) != 1:
        return False
    return all(pow(b, q // r, p) != 1 for r in sympy.factorint(q))


@pytest.mark.parametrize("it", PLANTED, ids=[it["tag"] for it in PLANTED])
def test_fixture_arithmetic(it):
    b, c, P0, P1, N0, N1, mod, res = it["item"]
    p, n_pl = it["p"], it["n_planted"]
    assert P0 <= p < P1 < (1 << 62) and P0 > 3
    t = (-c) % p
    assert pow(b, n_pl, p) == t
    if it.get("composite"):
        assert not sympy.isprime(p) and pow(2, p - 1, p) == 1
        assert min(sympy.factorint(p)) > 65521  # the device sieve keeps it
        return
    assert sympy.isprime(p)
    rel = it["relation"]
    if rel is None:
        return
    rp, M, q = rel
    assert rp == p and N0 <= M < N1 and (mod <= 1 or M % mod == res)
    assert pow(b, M, p) == t and _order_ok(b, p, q)
    # no smaller in-class solution: in-class solutions repeat with period lcm(q, mod)
    period = q * max(mod, 1) // math.gcd(q, max(mod, 1))
    assert M - period < N0


@pytest.mark.parametrize("it", PLANTED, ids=[it["tag"] for it in PLANTED])
def test_oracle_planted(oracle, it):
    item = oracle.make_item(*it["item"])
    if it.get("composite"):
        assert not oracle.is_prime64(it["p"])
        return
    assert oracle.is_prime64(it["p"])
    want = tuple(it["relation"]) if it["relation"] else None
    assert oracle.run_bsgs(it["p"], item) == want


def test_fixture_covers_both_fields_and_every_kind():
    tags = {it["tag"].split("/")[0] for it in PLANTED}
    assert {"c2", "c3", "f64top", "mont50", "c4", "c4cls", "top62", "lb12", "lb50", "lb18", "sm18"} <= tags
    assert any(it["p"] > (1 << 50) and it["relation"] for it in PLANTED)
    assert any(it["p"] > (1 << 61) and it["relation"] for it in PLANTED)
    assert any(it.get("composite") and 37 not in it["sprp_bases"] and 31 in it["sprp_bases"] for it in PLANTED)
