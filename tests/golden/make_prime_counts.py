"""Writes tests/golden/prime_counts.json: the exact number of primes in the P range of each
benchmark configuration, counted by the oracle's segmented sieve and, where sympy's prime
counting function is fast enough, checked against sympy.primepi."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402
import sympy  # noqa: E402

RANGES = [
    (10**9, 10**9 + 2**26),
    (10**12, 10**12 + 2**28),
    (95 * 10**12, 95 * 10**12 + 2**28),
    (4 * 10**18, 4 * 10**18 + 2**30),
    (10**10, 10**10 + 20_000_000),
]


def main():
    out = {}
    for lo, hi in RANGES:
        n = O.lib().or_sieve_primes(lo, hi, None, 0)
        if hi <= 10**13:
            assert n == sympy.primepi(hi - 1) - sympy.primepi(lo - 1), (lo, hi)
        out[f"{lo}:{hi}"] = int(n)
        print(lo, hi, n, flush=True)
    with open(os.path.join(HERE, "prime_counts.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
