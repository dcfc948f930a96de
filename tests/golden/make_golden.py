"""Writes tests/golden/relations_small.json: relations of small work items by brute force.

For every prime p of [P0, P1) and every in-class n of [N0, N1), p | b^n + c is tested directly
with pow(b, n, p); M is the smallest such n and q = sympy.n_order(b, p).  This shares no code
with the oracle or the product, so it pins both.  Primes below 1000 follow the reference's
fallback (run_bsgs_reference, Math.hpp:488-542): n is searched from 0 over [0, m*m) with
m = ceil(sqrt(min(p, N1))), ignoring N0 and the class.  The results.json object of each item is
restated here too (apply_relations Worker.hpp:527-559, Results.hpp:72-107) with compact JSON.
"""
import json
import math
import os

import sympy

HERE = os.path.dirname(os.path.abspath(__file__))

ITEMS = [
    # b, c, P0, P1, N0, N1, mod, res
    (2, -25, 1000, 40000, 3600, 9000, 2, 1),
    (2, -25, 1000, 20000, 3600, 9000, 6, 1),
    (2, -25, 1000, 20000, 3600, 9000, 1, 0),
    (10, 111, 1000, 30000, 0, 20000, 6, 4),
    (10, 111, 1000, 30000, 0, 20000, 1, 0),
    (3, -2, 5000, 40000, 100, 7000, 1, 0),
    (7, 3, 1001, 25000, 12345, 19999, 2, 0),
    (5, -4, 2000, 30000, 1, 30000, 4, 3),
    (6, 1, 1000, 20000, 50, 9050, 12, 5),
    (12, -7, 1500, 20000, 0, 6000, 3, 2),
    (2, 3, 100, 2000, 10, 400, 2, 1),      # primes below 1000: the reference's fallback
    (10, 111, 200, 1500, 0, 3000, 6, 4),
]


def isqrt_ceil(x):
    r = math.isqrt(x)
    return r if r * r == x else r + 1


def relation(p, b, c, N0, N1, mod, res):
    t = (-c) % p
    if p < 1000 or b % p == 0:
        if b % p == 0:
            return (p, N0, 1) if t == 0 else None
        m = isqrt_ceil(N1) if p > N1 else isqrt_ceil(p)
        for n in range(m * m):
            if pow(b, n, p) == t:
                return (p, n, int(sympy.n_order(b % p, p)))
        return None
    if t == 0:
        return None
    for n in range(max(N0, 0), N1):
        if mod > 1 and n % mod != res:
            continue
        if pow(b, n, p) == t:
            return (p, n, int(sympy.n_order(b % p, p)))
    return None


def results_object(item, rels, timestamp):
    b, c, P0, P1, N0, N1, mod, res = item
    count = N1 - N0
    factors = [[] for _ in range(count)]
    for p, M, q in rels:
        n = M + ((N0 - M + q - 1) // q) * q  # the first n >= N0 with n == M (mod q)
        while n < N1:
            # b^n + c == p is the prime itself, not a factored candidate (only while b^n <= 2p)
            small = n < 200 and b ** n <= 2 * p
            if not (small and b ** n + c == p):
                factors[n - N0].append(p)
            n += q
    prp, factored = [], []
    for i in range(count):
        n = N0 + i
        if mod > 1 and n % mod != res:
            continue
        if factors[i]:
            factored.append({"n": n, "factors": factors[i]})
        else:
            prp.append(n)
    obj = {"Timestamp": timestamp, "b": b, "c": c, "P0": P0, "P1": P1, "N0": N0, "N1": N1, "prp": prp,
           "factored": factored}
    return json.dumps(obj, separators=(",", ":"))


def main():
    out = []
    for item in ITEMS:
        b, c, P0, P1, N0, N1, mod, res = item
        rels = []
        for p in sympy.primerange(P0, P1):
            r = relation(p, b, c, N0, N1, mod, res)
            if r:
                rels.append(r)
        rec = {"item": list(item), "relations": [list(r) for r in rels]}
        if N1 - N0 <= 6000:  # results.json bytes for the short exponent ranges (keeps the file small)
            rec["results_json"] = results_object(item, rels, "2026-01-01 00:00:00 UTC+0")
        out.append(rec)
        print(item, len(rels), flush=True)
    with open(os.path.join(HERE, "relations_small.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
