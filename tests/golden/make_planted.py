"""Writes tests/golden/planted.json: work items with a relation planted at a known prime.

At large p a random range of P has almost no relations (about 0.06 per c3 item, 6e-5 per c4
item), so comparing relation lists there compares empty lists.  These items carry relations by
construction, at every magnitude the worker handles and in both of its fields:

* planted constant: pick a prime p and an exponent n, set c = -(b^n mod p); then p | b^n + c.
* large base: b of prime order d (about 1e6) modulo p, so that several n of the range solve and
  M must be the smallest of them, not the planted one.
* smooth p - 1 with the benchmark's own (b, c): the exponent is the discrete logarithm of -c,
  found by Pohlig-Hellman (sympy.discrete_log), and the N range is placed around it.
* strong pseudoprimes: composites with a planted solution that pass the base-2 Fermat test and
  Miller-Rabin to several bases; no relation may come out for them.

The expected relation of the planted prime is a closed form, shared with no product or oracle
code: with o = ord_p(b) (sympy.n_order, from sympy.factorint of p - 1), the solutions are
n == n_pl (mod o); M is the smallest of them with n >= N0 and n == res (mod mod), or none when
that is not below N1 (Math.hpp:377-486 defines M the same way); q = o.
"""
import json
import math
import os
import random

import sympy

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "planted.json")
W = 1 << 18  # P window of an item


def smallest_solution(p, b, n_pl, N0, N1, mod, res):
    """Smallest n >= N0, n == n_pl (mod ord_p(b)), n == res (mod mod), n < N1; or None."""
    o = int(sympy.n_order(b % p, p))
    mod = max(mod, 1)
    res %= mod
    g = math.gcd(o, mod)
    if (n_pl - res) % g:
        return None, o
    L = o // g * mod
    # n == n_pl (mod o) and n == res (mod mod): n = n_pl + o*k, o*k == res - n_pl (mod mod)
    k = ((res - n_pl) // g * pow(o // g, -1, mod // g)) % (mod // g) if mod // g > 1 else 0
    n0 = (n_pl + o * k) % L
    lo = max(N0, 0)
    n = lo + (n0 - lo) % L
    return (n if n < N1 else None), o


def window(p, where):
    """A P window of W integers holding p near its start, middle or end."""
    off = {"start": 5, "mid": W // 2 + 1, "end": W - 7}[where]
    P0 = p - off
    return P0, P0 + W


def planted_item(tag, p, b, n_pl, N0, N1, mod, res, where, fields):
    c = -pow(b, n_pl, p)
    P0, P1 = window(p, where)
    assert P0 > 3 and P1 < (1 << 62) and P0 <= p < P1
    M, o = smallest_solution(p, b, n_pl, N0, N1, mod, res)
    assert pow(b, n_pl, p) == (-c) % p
    return {"tag": tag, "item": [b, c, P0, P1, N0, N1, mod, res], "p": p, "n_planted": n_pl,
            "relation": [p, M, o] if M is not None else None, "fields": fields}


def last_in_class(N1, mod, res):
    n = N1 - 1
    return n - ((n - res) % mod)


def first_in_class(N0, mod, res):
    return N0 + ((res - N0) % mod)


def fields_of(p):
    """Fields an item can run in: FP64 only below 2^50 (else Montgomery); Montgomery always."""
    return [0, 1] if p + W < (1 << 50) else [0]


def planted_constant_items():
    items = []
    cases = [
        # tag, prime near, b, N0, N1, mod, res
        ("c2", 10**12 + 123456789, 10, 0, 10_000_000, 6, 4),
        ("c3", 95 * 10**12 + 987654321, 2, 3_600_000, 5_000_000, 2, 1),
        ("f64top", (1 << 50) - W // 2, 3, 0, 10_000_000, 1, 0),
        ("mont50", (1 << 50) + 3 * W, 7, 0, 10_000_000, 2, 0),
        ("c4", 4 * 10**18 + 5 * 10**8, 3, 0, 10_000_000, 1, 0),
        ("c4cls", 4 * 10**18 + 7 * 10**8, 3, 0, 10_000_000, 6, 5),
        ("top62", (1 << 62) - 3 * W, 5, 0, 10_000_000, 6, 1),
    ]
    for tag, near, b, N0, N1, mod, res in cases:
        p = int(sympy.nextprime(near))
        first = first_in_class(N0, mod, res)
        last = last_in_class(N1, mod, res)
        mid = first_in_class((N0 + N1) // 2, mod, res)
        # first: with N0 = 0, b^first < p, so c = -b^first and every prime divides b^first + c = 0
        # (a degenerate item: M = first for every prime of the window); near: the first in-class n
        # past 64, where b^n > p
        near = first_in_class(max(N0, 64), mod, res)
        cands = [(first, "start"), (mid, "mid"), (last, "end")]
        if near != first:
            cands.insert(1, (near, "end"))
        for n_pl, where in cands:
            items.append(planted_item(f"{tag}/n{n_pl}", p, b, n_pl, N0, N1, mod, res, where, fields_of(p)))
        if mod > 1:  # an exponent out of class: no in-class n below N1 solves
            n_out = mid + 1
            it = planted_item(f"{tag}/out{n_out}", p, b, n_out, N0, N1, mod, res, "mid", fields_of(p))
            assert it["relation"] is None
            items.append(it)
        # just below N0 and at N1: outside the range
        if N0 > 0:
            it = planted_item(f"{tag}/below{N0 - 1}", p, b, N0 - 1, N0, N1, mod, 0 if mod == 1 else (N0 - 1) % mod,
                              "mid", fields_of(p))
            assert it["relation"] is None
            items.append(it)
        it = planted_item(f"{tag}/at{N1}", p, b, N1, N0, N1, 1, 0, "end", fields_of(p))
        assert it["relation"] is None
        items.append(it)
    return items


def large_base_items(rng):
    """b of prime order d modulo p, d about 1e6: 5 to 10 solutions in [0, 1e7); M is the smallest."""
    items = []
    for tag, near, mod, res in (("lb12", 10**12, 1, 0), ("lb50", (1 << 50) + 10**9, 1, 0),
                                ("lb18", 4 * 10**18, 1, 0), ("lb18c2", 4 * 10**18 + 10**12, 2, 1)):
        d = int(sympy.nextprime(rng.randrange(1_000_000, 2_000_000)))
        k = near // d + 1
        while not sympy.isprime(k * d + 1):
            k += 1
        p = k * d + 1
        h = 2
        while pow(h, (p - 1) // d, p) == 1:
            h += 1
        b = pow(h, (p - 1) // d, p)
        assert int(sympy.n_order(b, p)) == d and b < (1 << 63)
        N0, N1 = 0, 10_000_000
        n_pl = last_in_class(N1, mod, res)
        it = planted_item(f"{tag}/d{d}", p, b, n_pl, N0, N1, mod, res, "mid", fields_of(p))
        M = it["relation"][1]
        assert M < n_pl and (n_pl - M) % d == 0 and M < d * mod
        # brute force: no in-class n in [N0, M) solves, and M does
        t, x = (-it["item"][1]) % p, pow(b, N0, p)
        for n in range(N0, M + 1):
            if n % mod == res:
                assert (x == t) == (n == M)
            x = x * b % p
        items.append(it)
    return items


def smooth_prime(rng, near):
    """A prime p in [near, near * (1 + 2^-12)) with p - 1 a product of primes below 2^21."""
    while True:
        s = 2
        while s < near >> 44:
            s *= int(sympy.prevprime(rng.randrange(1 << 10, 1 << 20)))
        if s > near >> 24:
            continue
        k = near // s  # at least 2^24: a step of s moves p by at most near * 2^-24
        for kk in range(k, k + 4096):
            p = s * kk + 1
            if p >= near and max(sympy.factorint(kk) or {1: 1}) < (1 << 21) and sympy.isprime(p):
                return p


def smooth_items(rng):
    """The benchmark's own (b, c) at a prime where the discrete logarithm of -c is cheap: the N range
    is placed around it (N about p, far past the benchmark's N ranges)."""
    items = []
    for tag, near, b, c, mod in (("sm18", 4 * 10**18, 3, -2, 1), ("sm12", 10**12, 10, 111, 6),
                                 ("sm13", 95 * 10**12, 2, -25, 2)):
        while True:
            p = smooth_prime(rng, near)
            try:
                n = int(sympy.discrete_log(p, (-c) % p, b))
            except ValueError:
                near = p + 2
                continue
            if n > 10_000_000:
                break
            near = p + 2
        assert p < (1 << 62) - W
        assert pow(b, n, p) == (-c) % p
        res = n % mod
        N0 = n - 2_345_678
        N1 = N0 + 10_000_000
        P0, P1 = window(p, "mid")
        M, o = smallest_solution(p, b, n, N0, N1, mod, res)
        assert M is not None and M <= n
        items.append({"tag": f"{tag}/p{p}", "item": [b, c, P0, P1, N0, N1, mod, res], "p": p, "n_planted": n,
                      "relation": [p, M, o], "fields": fields_of(p)})
    return items


def is_sprp(n, a):
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(a, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


MR_BASES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]


def strong_pseudoprimes():
    """Composites with no prime factor <= 65521 (so the device sieve keeps them) that are strong
    probable primes to a prefix of the Miller-Rabin bases: 3825123056546413051 to 2..31 (only 37
    proves it composite), 341550071728321 to 2..17, and numbers p * (2p - 1) found by search that
    pass base 2, or 2 and 3."""
    out = [3825123056546413051, 341550071728321]
    want_prefix = {1: 2, 2: 2}
    found = {}
    p = 65537
    while len(found) < sum(want_prefix.values()) and p < 50_000_000:
        p = int(sympy.nextprime(p))
        q = 2 * p - 1
        if not sympy.isprime(q):
            continue
        n = p * q
        k = 0
        while k < len(MR_BASES) and is_sprp(n, MR_BASES[k]):
            k += 1
        if k in want_prefix and sum(1 for v in found.values() if v == k) < want_prefix[k]:
            found[n] = k
    out += sorted(found)
    res = []
    for n in out:
        f = sympy.factorint(n)
        assert len(f) > 1 or max(f.values()) > 1
        assert min(f) > 65521 and pow(2, n - 1, n) == 1
        k = 0
        while k < len(MR_BASES) and is_sprp(n, MR_BASES[k]):
            k += 1
        res.append((n, k))
    return res


def order_mod_composite(b, n):
    """ord_n(b) from the factorisation of n (lcm over the prime powers)."""
    o = 1
    for r, e in sympy.factorint(n).items():
        o = math.lcm(o, int(sympy.n_order(b % r, r)) * r ** (e - 1))
    assert pow(b, o, n) == 1
    return o


def pseudoprime_items():
    """A planted solution at each strong pseudoprime.  The N range is shorter than ord_n(b) (about
    4e3 to 3e7 here), so the planted solution is the only one in range: then no short cycle or hit
    list overflow leaves the number to the host in plain mode, the device writes a record for it,
    and only its Miller-Rabin test can drop it."""
    items = []
    for n, k in strong_pseudoprimes():
        for b in (2, 3):
            if not is_sprp(n, b) and b == 3:
                continue  # base 3 already proves it composite: one item (b = 2) is enough
            o = order_mod_composite(b, n)
            N0 = 4_000_000
            N1 = N0 + min(100_000, o // 2)
            n_pl = (N0 + N1) // 2 + 1
            c = -pow(b, n_pl, n)
            P0, P1 = window(n, "mid")
            items.append({"tag": f"spsp{k}/{n}/b{b}", "item": [b, c, P0, P1, N0, N1, 1, 0], "p": n,
                          "n_planted": n_pl, "relation": None, "composite": True, "sprp_bases": MR_BASES[:k],
                          "order": o, "fields": fields_of(n)})
    return items


# Small b and c in the Montgomery field (p > 2^50): primes p of b^n + c found by factoring b^n + c
# (sympy.factorint, n up to 75), so that b and t = -c stay small integers, as in c4's (3, -2), and
# the order-mode setup takes its small-base powers (windows of 4 bits for |b|, |c| <= 3, 3 bits up
# to 10).  c > 0 makes t negative.
SMALL_BASE_CASES = [
    (3, -2, 37, 450283905890997361),
    (3, -2, 40, 11246684050931479),
    (3, 2, 36, 150094635296999123),
    (3, 2, 41, 1387597351233433),
    (2, 3, 61, 461168601842738791),
    (7, -5, 36, 2947742064732341939),
]


def small_base_items():
    items = []
    for k, (b, c, n_pl, p) in enumerate(SMALL_BASE_CASES):
        assert sympy.isprime(p) and (pow(b, n_pl, p) + c) % p == 0 and p > (1 << 50)
        mod, res = (6, n_pl % 6) if k % 2 else (1, 0)
        P0, P1 = window(p, ("start", "mid", "end")[k % 3])
        M, o = smallest_solution(p, b, n_pl, 0, 10**7, mod, res)
        items.append({"tag": f"mont-small/{b},{c},n{n_pl}", "item": [b, c, P0, P1, 0, 10**7, mod, res], "p": p,
                      "n_planted": n_pl, "relation": [p, M, o] if M is not None else None, "fields": fields_of(p)})
    return items


def main():
    rng = random.Random(20261204)
    items = (planted_constant_items() + large_base_items(rng) + smooth_items(rng) + pseudoprime_items() +
             small_base_items())
    with open(OUT, "w") as f:
        json.dump({"window": W, "items": items}, f, indent=1)
        f.write("\n")
    print(f"{len(items)} items -> {OUT}")


if __name__ == "__main__":
    main()
