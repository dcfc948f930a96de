"""The served workload: a seeded stream of Factor= work items with Poisson arrivals.

Each item is drawn from numpy.random.default_rng(seed), field by field in this order:
  gap    exponential with mean 1 (unit-rate arrivals; a caller scales the gaps by 1/rate)
  b      uniform over {2, 3, 5, 6, 7, 10, 12}
  c      uniform over [-9999, 9999] without 0
  P0     floor(10^U(9, 15))
  width  floor(2^U(24, 30))                    (P1 = P0 + width)
  N0     floor(U(0, 5e6))
  Nw     floor(10^U(5, 7))                     (N1 = N0 + Nw)
  mod    uniform over {none, 2, 6}; res uniform in [0, mod)
and c is then drawn again (same generator) until b^n + c is not identically reducible:
gcd(b, c) = 1, no fixed divisor (the gcd of b^n + c over n >= 0 is 1), not -c a square when
every in-class n is even, and not +-c a cube when every in-class n is a multiple of 3.

The item fields are those of a worktodo.txt line, Factor=b,c,P0,P1,N0,N1[,mod,res] (the
reference's parse_line, Types.hpp:14-26 and Worktodo.hpp).
"""
import math

import numpy as np

BASES = (2, 3, 5, 6, 7, 10, 12)
MODS = (1, 2, 6)  # 1: no class restriction


def _is_power(v, k):
    """v == s^k for an integer s >= 1."""
    if v <= 0:
        return False
    s = round(v ** (1.0 / k))
    return any((s + d) ** k == v for d in (-1, 0, 1) if s + d >= 1)


def fixed_divisor(b, c):
    """gcd over n >= 0 of b^n + c (it is reached within the first terms: b^n + c - (1 + c) is a
    multiple of b - 1)."""
    g = 0
    for n in range(0, 24):
        g = math.gcd(g, abs(b ** n + c))
    return g


def reducible(b, c, mod, res):
    """b^n + c factors for every in-class n (see the module docstring)."""
    if math.gcd(b, abs(c)) != 1 or fixed_divisor(b, c) != 1:
        return True
    every = math.gcd(mod, res) if mod > 1 else 1  # every in-class n is a multiple of this
    if every % 2 == 0 and _is_power(-c, 2):
        return True
    if every % 3 == 0 and (_is_power(c, 3) or _is_power(-c, 3)):
        return True
    return False


def serving_mix(seed=0, n=256):
    """n items as dicts (b, c, P0, P1, N0, N1, mod, res, gap) of the seeded stream."""
    rng = np.random.default_rng(seed)

    def draw_c():
        while True:
            c = int(rng.integers(-9999, 10000))
            if c != 0:
                return c

    out = []
    for _ in range(n):
        gap = float(rng.exponential(1.0))
        b = int(BASES[rng.integers(len(BASES))])
        c = draw_c()
        P0 = int(math.floor(10.0 ** rng.uniform(9, 15)))
        width = int(math.floor(2.0 ** rng.uniform(24, 30)))
        N0 = int(math.floor(rng.uniform(0, 5e6)))
        nwidth = int(math.floor(10.0 ** rng.uniform(5, 7)))
        mod = int(MODS[rng.integers(len(MODS))])
        res = int(rng.integers(mod)) if mod > 1 else 0
        while reducible(b, c, mod, res):
            c = draw_c()
        out.append(dict(b=b, c=c, P0=P0, P1=P0 + width, N0=N0, N1=N0 + nwidth, mod=mod, res=res, gap=gap))
    return out


def worktodo_line(it):
    """The worktodo.txt line of an item."""
    s = f"Factor={it['b']},{it['c']},{it['P0']},{it['P1']},{it['N0']},{it['N1']}"
    if it["mod"] > 1:
        s += f",{it['mod']},{it['res']}"
    return s


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="Prints the seeded serving mix as worktodo.txt lines.")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("-n", type=int, default=256)
    args = ap.parse_args()
    for it in serving_mix(args.seed, args.n):
        print(worktodo_line(it))
