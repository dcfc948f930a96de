"""The drop-in exactly as the reference's host drives it (tests/dropin_check.cpp): the unmodified
adapter include/mps_gpu_siever.hpp behind the reference's WorkItem, GPU_threads = 2 feeders on one
handle, 2^26-integer chunks, and the host's own sort / dedup / is_prime64 / run_bsgs after
sieve_filter (Worker.hpp:216-228) or after filter (Worker.hpp:245-257).  Its relations must be
the oracle's CPU path's.  The planted items carry relations at p ~ 4e18 (Montgomery field), and
an item with a relation at every prime makes sieve_filter and filter return more numbers than the
adapter's first buffer holds, so its MPS_CAPACITY retry runs."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "_bin", "dropin_check")
PLANTED = {it["tag"]: it for it in json.load(open(os.path.join(HERE, "golden", "planted.json")))["items"]}

C1 = (2, -25, 10**9, 10**9 + 2**26, 3_600_000, 5_000_000, 2, 1)


def run_dropin(mode, item, chunk=1 << 26, threads=2):
    assert os.path.exists(EXE), "tests/_bin/dropin_check is not built (run __graft_entry__.build())"
    args = [EXE, mode] + [str(v) for v in item] + [str(chunk), str(threads)]
    print(" ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    rels = []
    for line in r.stdout.splitlines():
        p, M, q = (int(x) for x in line.split(","))
        rels.append((p, M, q))
    return rels


def oracle_rels(oracle, item):
    b, c, P0, P1, N0, N1, mod, res = item
    want, _ = oracle.process_range(oracle.make_item(*item), P0, P1, 1 << 22, 16)
    return want


@pytest.mark.parametrize("mode", ["sieve", "filter"])
def test_dropin_c1(oracle, mode):
    want = oracle_rels(oracle, C1)
    assert len(want) == 2224
    assert run_dropin(mode, C1) == want


@pytest.mark.parametrize("tag", ["c4/n5000000", "c4cls/n9999995", "lb18/d1008223", "top62/n9999997", "sm18"])
@pytest.mark.parametrize("mode", ["sieve", "filter"])
def test_dropin_planted_high_p(oracle, mode, tag):
    it = next(v for k, v in PLANTED.items() if k == tag or k.startswith(tag + "/"))
    item = tuple(it["item"])
    want = oracle_rels(oracle, item)
    assert tuple(it["relation"]) in want
    assert run_dropin(mode, item, chunk=1 << 17) == want


@pytest.mark.parametrize("mode", ["sieve", "filter"])
def test_dropin_capacity_retry(oracle, mode):
    """c = -10^4: every prime of the window divides 10^4 + c = 0, so all of them come back."""
    item = tuple(PLANTED["c2/n4"]["item"])
    want = oracle_rels(oracle, item)
    assert len(want) > 4096 + 1024  # past both of the adapter's first buffers (one 2^18 chunk)
    assert run_dropin(mode, item, chunk=1 << 18, threads=1) == want
