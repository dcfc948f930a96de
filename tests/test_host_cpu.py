"""CPU checks of the product's host side (no GPU needed): libmps.so loads and exports every
symbol the C headers declare, and the exact host checks (is_prime64, run_bsgs,
multiplicative_order) and the results.json writer agree with the oracle and the golden fixtures."""
import json
import os
import random
import re

import pytest
import sympy

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

with open(os.path.join(HERE, "golden", "relations_small.json")) as f:
    GOLDEN = json.load(f)


def declared_functions():
    names = set()
    for h in ("mps_gpu.h", "mps_host.h", "mps_batch.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mps_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol(mps):
    L = mps.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    assert set(mps.EXPORTS) <= names


def test_create_without_gpu_fails_cleanly(mps):
    """GpuSiever::init without a usable device: no handle and a 'GPU disabled' reason
    (Gpu.hpp:1054), never a crash."""
    import ctypes
    g = mps.GpuSiever()
    if g.init(0):  # a GPU is present: nothing to check here
        g.close()
        pytest.skip("a HIP device is present")
    assert g.banner.startswith("GPU disabled")


def test_host_is_prime64(mps, oracle):
    rng = random.Random(5)
    for _ in range(3000):
        n = rng.randrange(2 ** rng.randrange(2, 64))
        assert mps.is_prime64(n) == oracle.is_prime64(n) == sympy.isprime(n)
    assert not mps.is_prime64(3825123056546413051)


def test_host_multiplicative_order(mps, oracle):
    rng = random.Random(6)
    for bits in (20, 40, 62):
        for _ in range(30):
            p = sympy.randprime(2 ** (bits - 1), 2 ** bits)
            b = rng.randrange(2, 1000)
            assert mps.multiplicative_order(b, p) == oracle.multiplicative_order(b, p)


@pytest.mark.parametrize("rec", GOLDEN[:6], ids=lambda r: ",".join(map(str, r["item"])))
def test_host_run_bsgs_matches_golden(mps, rec):
    b, c, P0, P1, N0, N1, mod, res = rec["item"]
    item = mps.work_item(b, c, P0, P1, N0, N1, mod, res)
    want = {r[0]: tuple(r) for r in rec["relations"]}
    for p in sympy.primerange(P0, P1):
        assert mps.run_bsgs(p, item) == want.get(p), p


@pytest.mark.parametrize("rec", [r for r in GOLDEN if "results_json" in r], ids=lambda r: ",".join(map(str, r["item"])))
def test_host_results_json_matches_golden(mps, rec):
    b, c, P0, P1, N0, N1, mod, res = rec["item"]
    item = mps.work_item(b, c, P0, P1, N0, N1, mod, res)
    rels = [tuple(r) for r in rec["relations"]]
    assert mps.results_json(item, rels, "2026-01-01 00:00:00 UTC+0") == rec["results_json"]


def test_host_results_json_matches_oracle_large(mps, oracle):
    """A long exponent range and relations with small q (many factored n)."""
    item = mps.work_item(10, 111, 10 ** 6, 10 ** 6 + 50000, 0, 200000, 6, 4)
    oit = oracle.make_item(10, 111, 10 ** 6, 10 ** 6 + 50000, 0, 200000, 6, 4)
    rels, _ = oracle.process_range(oit, 10 ** 6, 10 ** 6 + 50000, 1 << 20, 4)
    assert mps.results_json(item, rels, "t") == oracle.results_json(oit, rels, "t")


def test_cli_without_gpu_fails_loudly(tmp_path, mps):
    """No CPU sieve path: without a usable GPU the command-line siever stops with exit 1."""
    import subprocess
    g = mps.GpuSiever()
    if g.init(0):
        g.close()
        pytest.skip("a HIP device is present")
    binary = os.path.join(ROOT, "mass-prime-siever_amd", "mps_siever")
    (tmp_path / "worktodo.txt").write_text("Factor=2,-25,1000000000,1001000000,3600000,5000000,2,1\n")
    (tmp_path / "prime.ini").write_text("CPU_count=0\nMemoryAllowed=64\nUseGPU=1\n")
    r = subprocess.run([binary], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "GPU disabled" in (tmp_path / "siever.log").read_text()


def test_dropin_adapter_compiles_and_declines_without_gpu(tmp_path):
    """include/mps_gpu_siever.hpp, the adapter a reference host includes in place of Gpu.hpp,
    compiles with -Wall -Werror against a WorkItem with 128-bit fields (Types.hpp:14-26) and links
    against libmps.so; without a usable device it reports 'GPU disabled' and stops (GPU_strict=1)."""
    import subprocess
    exe = str(tmp_path / "dropin_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(HERE, "dropin_check.cpp"), "-o", exe,
                    "-L" + os.path.join(ROOT, "mass-prime-siever_amd"), "-lmps", "-L" + os.path.join(ROOT, "oracle"),
                    "-loracle", "-Wl,-rpath," + os.path.join(ROOT, "mass-prime-siever_amd"),
                    "-Wl,-rpath," + os.path.join(ROOT, "oracle"), "-lpthread"], check=True)
    r = subprocess.run([exe, "sieve", "2", "-25", "1000000000", "1001048576", "3600000", "5000000", "2", "1"],
                       capture_output=True, text=True, timeout=120)
    if "GPU enabled" in r.stderr:
        pytest.skip("a HIP device is present")
    assert r.returncode == 1 and "GPU disabled" in r.stderr
