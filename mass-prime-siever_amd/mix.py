"""The served workload: a seeded stream of Factor= work items with Poisson arrivals.

Each item is drawn as follows (seeded, so a seed names one stream):
  b      uniform from {2, 3, 5, 6, 7, 10, 12}
  c      uniform nonzero integer in [-9999, 9999], odd or even, with gcd(b, c) = 1 and
         b^n + c not identically reducible on the item's exponent class (no fixed prime factor
         below 200 and no algebraic factorisation: c = -s^2 when every n is even, c = +-s^3 when
         every n is a multiple of 3)
  P0     log-uniform in [1e9, 1e15]
  P1-P0  log-uniform in [2^24, 2^30]
  N0     uniform in [0, 5e6]
  N1-N0  log-uniform in [1e5, 1e7]
  class  mod from {none, 2, 6} with res uniform in [0, mod)
Arrival gaps are exponential with mean 1 (unit rate); a caller scales them by 1/rate.

The item fields are those of a worktodo.txt line, Factor=b,c,P0,P1,N0,N1[,mod,res] (the
reference's parse_line, Types.hpp:14-26 and Worktodo.hpp).
"""
import math
import random

BASES = (2, 3, 5, 6, 7, 10, 12)
MODS = (1, 2, 6)


def _fixed_prime_factor(b, c, n0, mod):
    """A prime r < 200 dividing b^n + c for every n == n0 (mod mod), n >= n0, or 0."""
    for r in range(2, 200):
        if any(r % d == 0 for d in range(2, int(math.isqrt(r)) + 1)):
            continue
        # b^n mod r has period dividing r - 1; stepping n by mod covers a full orbit in r steps
        if all((pow(b, n0 + mod * k, r) + c) % r == 0 for k in range(r)):
            return r
    return 0


def _is_power(v, k):
    """v == s^k for an integer s >= 1."""
    if v <= 0:
        return False
    s = round(v ** (1.0 / k))
    return any((s + d) ** k == v for d in (-1, 0, 1) if s + d >= 1)


def reducible(b, c, mod, res, n0):
    """b^n + c factors for every n of the class (see the module docstring)."""
    if math.gcd(b, abs(c)) != 1:
        return True
    first = n0 + ((res - n0) % mod) if mod > 1 else n0
    if _fixed_prime_factor(b, c, first, mod):
        return True
    every = math.gcd(mod, res) if mod > 1 else 1  # every n of the class is a multiple of this
    if every % 2 == 0 and _is_power(-c, 2):
        return True
    if every % 3 == 0 and (_is_power(c, 3) or _is_power(-c, 3)):
        return True
    return False


def serving_mix(seed=0, n=256):
    """n items as dicts (b, c, P0, P1, N0, N1, mod, res, gap) of the seeded stream."""
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        b = rng.choice(BASES)
        P0 = int(math.exp(rng.uniform(math.log(1e9), math.log(1e15))))
        width = int(2.0 ** rng.uniform(24, 30))
        N0 = rng.randint(0, 5_000_000)
        nwidth = int(10.0 ** rng.uniform(5, 7))
        mod = rng.choice(MODS)
        res = rng.randrange(mod) if mod > 1 else 0
        while True:
            c = rng.randint(-9999, 9999)
            if c != 0 and not reducible(b, c, mod, res, N0):
                break
        gap = rng.expovariate(1.0)
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
