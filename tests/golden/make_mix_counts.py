"""Writes tests/golden/mix_seed0.json: the 256 items of the seeded serving mix (seed 0) and the
exact number of primes in each item's P range, counted by the oracle's segmented sieve (whose
counts are checked against sympy.primepi in make_prime_counts.py)."""
import json
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "mass-prime-siever_amd"))
import mix  # noqa: E402
import oracle as O  # noqa: E402


def count(rng):
    lo, hi = rng
    return int(O.lib().or_sieve_primes(lo, hi, None, 0))


def main():
    items = mix.serving_mix(seed=0, n=256)
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        counts = pool.map(count, [(it["P0"], it["P1"]) for it in items], chunksize=1)
    for it, n in zip(items, counts):
        it["primes"] = n
    with open(os.path.join(HERE, "mix_seed0.json"), "w") as f:
        json.dump({"seed": 0, "items": items}, f, indent=0)
        f.write("\n")
    print(len(items), "items,", sum(counts), "primes")


if __name__ == "__main__":
    main()
