"""Seeded random work items, GPU worker against the oracle, in every setup mode.

Fifty-two items drawn from random.Random(20261204): bases {2, 3, 5, 6, 7, 10, 12}, class moduli
{1, 2, 3, 4, 5, 6, 7, 10, 12, 15, 30} with a random residue, N ranges of 10^3 to 10^6 exponents from
N0 below 2*10^6, and P windows of 2^16 integers:
  * 24 with a small constant c (|c| < 10^4) and P0 log-uniform in [10^4, 10^10], where relations
    are common;
  * 28 with a relation planted at a prime p of the window (c = -(b^n mod p) for an in-range n),
    P0 log-uniform in [10^8, 4*10^18] (24 of them) or uniform in [10^18, 4*10^18] (4), so both
    device fields and every magnitude up to the 62-bit limit carry relations.
Each item runs with classes off (plain mode) and with classes on under class policies 0, 1 and 2
(class and order mode), and below 2^50 also with the Montgomery field forced.  The relations must
equal the oracle's CPU path over the same window (sorted p, M, q; Math.hpp:377-486 defines M), and
across the whole set most items must carry relations, so the comparison is not of empty lists.
The odd class moduli exercise the filter's character rule where the class has one character of
b^n only when (b/p) = 1."""
import math
import random

import pytest

BASES = (2, 3, 5, 6, 7, 10, 12)
MODS = (1, 2, 3, 4, 5, 6, 7, 10, 12, 15, 30)
W = 1 << 16


def _log_uniform(rng, lo, hi):
    return int(math.exp(rng.uniform(math.log(lo), math.log(hi))))


def _n_range(rng):
    N0 = rng.randrange(0, 2_000_000)
    return N0, N0 + _log_uniform(rng, 1_000, 1_000_000)


def _class(rng):
    mod = rng.choice(MODS)
    return mod, (rng.randrange(mod) if mod > 1 else 0)


def _items(is_prime):
    rng = random.Random(20261204)
    out = []
    for k in range(24):
        b = rng.choice(BASES)
        c = rng.choice([-1, 1]) * rng.randrange(1, 10_000)
        P0 = _log_uniform(rng, 10**4, 10**10)
        N0, N1 = _n_range(rng)
        mod, res = _class(rng)
        out.append((f"small{k}", (b, c, P0, P0 + W, N0, N1, mod, res)))
    for k in range(28):
        b = rng.choice(BASES)
        P0 = _log_uniform(rng, 10**8, 4 * 10**18) if k < 24 else rng.randrange(10**18, 4 * 10**18)
        N0, N1 = _n_range(rng)
        mod, res = _class(rng)
        p = P0 + rng.randrange(W // 8, W - W // 8) | 1
        while not is_prime(p):
            p += 2
        assert p < P0 + W and p % b != 0
        # an in-class exponent of the range
        n = N0 + (res - N0) % mod + mod * rng.randrange(0, max(1, (N1 - N0) // mod))
        if n >= N1:
            n = N0 + (res - N0) % mod
        c = -pow(b, n, p)
        out.append((f"planted{k}", (b, c, P0, P0 + W, N0, N1, mod, res)))
    return out


MODES = {"plain": (False, 1, -1), "policy0": (True, 0, -1), "policy1": (True, 1, -1), "policy2": (True, 2, -1),
         "policy2-mont": (True, 2, 0)}

_cache = {}


def _setup(oracle):
    if "items" not in _cache:
        _cache["items"] = _items(lambda x: bool(oracle.is_prime64(x)))
    return _cache["items"]


def _want(oracle, item):
    if item not in _cache:
        _cache[item], _ = oracle.process_range(oracle.make_item(*item), item[2], item[3], 1 << 16, 8)
    return _cache[item]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("k", range(52))
def test_random_item(gpu, mps, oracle, k, mode):
    tag, item = _setup(oracle)[k]
    classes, policy, field = MODES[mode]
    if field == 0 and item[3] > (1 << 50):
        pytest.skip("the Montgomery field is the default above 2^50")
    print(f"random item {tag} {mode}: {item}", flush=True)
    want = _want(oracle, item)
    gpu.set_classes(classes)
    gpu.set_class_policy(policy)
    gpu.configure(field, 0)
    try:
        rels, st = gpu.run_item(mps.work_item(*item), chunk_ints=1 << 16)
    finally:
        gpu.set_classes(True)
        gpu.set_class_policy(1)
        gpu.configure(-1, 0)
    assert st.declined == 0
    assert rels == want


def test_random_items_carry_relations(oracle):
    """(CPU) Most items of the set have relations (the planted ones all do), so the parity above is
    not a comparison of empty lists."""
    items = _setup(oracle)
    with_rel = [tag for tag, it in items if _want(oracle, it)]
    planted = [tag for tag, _ in items if tag.startswith("planted")]
    assert set(planted) <= set(with_rel)
    assert len(with_rel) >= 40
