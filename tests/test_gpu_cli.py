"""The command-line siever (mass-prime-siever_amd/mps_siever) on the GPU, compared with the
oracle's CPU path the way the reference's own self-test compares its GPU and CPU runs
(kaggle/kaggle_run.py:64-95): sorted relations.p lines must be equal, with at least one relation;
the log must say "GPU enabled"; and results.json must hold the oracle's object for the item."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mass-prime-siever_amd", "mps_siever")


def run_siever(tmp_path, lines, ini):
    (tmp_path / "worktodo.txt").write_text("".join(l + "\n" for l in lines))
    (tmp_path / "prime.ini").write_text(ini)
    r = subprocess.run([BIN], cwd=tmp_path, stdin=subprocess.DEVNULL, capture_output=True, text=True, timeout=300)
    return r


def test_cli_self_test_item(tmp_path, oracle):
    item = (2, -25, 10_000_000_000, 10_020_000_000, 3_600_000, 5_000_000, 2, 1)
    r = run_siever(tmp_path, ["Factor=" + ",".join(map(str, item))],
                   "CPU_count=0\nMemoryAllowed=16\nUseGPU=1\nGPU_device=0\nGPU_threads=2\nGPU_sieve=1\nGPU_strict=1\n")
    print(r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    log = (tmp_path / "siever.log").read_text()
    assert "GPU enabled" in log
    got = sorted((tmp_path / "relations.p").read_text().split("\n"))
    oit = oracle.make_item(*item)
    rels, _ = oracle.process_range(oit, item[2], item[3], 1 << 22, 8)
    want = sorted([f"{p}, {M}, {q}" for p, M, q in rels] + [""])
    assert got == want
    assert len(want) > 1
    res = json.loads((tmp_path / "results.json").read_text())
    assert len(res) == 1
    ours = res[0]
    ref = json.loads(oracle.results_json(oit, rels, ours["Timestamp"]))
    assert ours == ref


def test_cli_two_items_append_results(tmp_path, oracle):
    items = [(10, 111, 10 ** 9, 10 ** 9 + 2 ** 20, 0, 200_000, 6, 4), (3, -2, 10 ** 9, 10 ** 9 + 2 ** 19, 0, 100_000)]
    r = run_siever(tmp_path, ["Factor=" + ",".join(map(str, it)) for it in items],
                   "CPU_count=0\nMemoryAllowed=64\nUseGPU=1\nGPU_sieve=1\nGPU_classes=1\n")
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads((tmp_path / "results.json").read_text())
    assert len(res) == 2
    for it, obj in zip(items, res):
        full = it + (1, 0) if len(it) == 6 else it
        oit = oracle.make_item(*full)
        rels, _ = oracle.process_range(oit, full[2], full[3], 1 << 20, 8)
        assert obj["prp"] == json.loads(oracle.results_json(oit, rels, "t"))["prp"]


@pytest.mark.parametrize("mode", ["GPU_batch=1\n", "GPU_autotune=1\n", "GPU_batch=1\nGPU_autotune=1\n"])
def test_cli_batch_and_autotune_modes_match_sequential(tmp_path, mode):
    """GPU_batch=1 (every item through the batch engine) and GPU_autotune=1 give the same
    relations.p and results.json objects, in worktodo order, as the plain sequential run."""
    items = [(10, 111, 10 ** 9, 10 ** 9 + 2 ** 21, 0, 2_000_000, 6, 4), (3, -2, 10 ** 9, 10 ** 9 + 2 ** 20, 0, 500_000),
             (2, -25, 10 ** 10, 10 ** 10 + 2 ** 21, 3_600_000, 5_000_000, 2, 1)]
    lines = ["Factor=" + ",".join(map(str, it)) for it in items]
    base = "CPU_count=0\nMemoryAllowed=8\nUseGPU=1\nGPU_sieve=1\nGPU_classes=1\n"
    seq, alt = tmp_path / "seq", tmp_path / "alt"
    seq.mkdir()
    alt.mkdir()
    print(f"cli mode {mode!r}", flush=True)
    r1 = run_siever(seq, lines, base)
    r2 = run_siever(alt, lines, base + mode)
    assert r1.returncode == 0 and r2.returncode == 0, (r1.stderr[-1000:], r2.stderr[-1000:])
    assert (seq / "relations.p").read_text() == (alt / "relations.p").read_text()
    a = json.loads((seq / "results.json").read_text())
    b = json.loads((alt / "results.json").read_text())
    assert len(a) == len(b) == 3
    for x, y in zip(a, b):
        x.pop("Timestamp", None)
        y.pop("Timestamp", None)
        assert x == y


def test_cli_resumes_from_core_files(tmp_path, oracle):
    """read_checkpoint (Worker.hpp:369-463): the run resumes from the smallest last-line p of the
    core files, drops the lines at or above it, and the merged relations equal a full run's."""
    item = (10, 111, 10 ** 9, 10 ** 9 + 2 ** 22, 0, 2_000_000, 6, 4)
    oit = oracle.make_item(*item)
    full, _ = oracle.process_range(oit, item[2], item[3], 1 << 20, 8)
    assert len(full) > 8
    k = len(full) // 2
    lines = [f"{p}, {M}, {q}" for p, M, q in full]
    (tmp_path / "core_0.p").write_text("\n".join(lines[:k + 1]) + "\n")         # last p = full[k]
    (tmp_path / "core_1.p").write_text("\n".join(lines[k + 1:k + 4]) + "\n")    # later p: trimmed away
    r = run_siever(tmp_path, ["Factor=" + ",".join(map(str, item))],
                   "CPU_count=0\nMemoryAllowed=1\nUseGPU=1\nGPU_sieve=1\nGPU_classes=1\n")
    assert r.returncode == 0, r.stderr[-2000:]
    log = (tmp_path / "siever.log").read_text()
    assert f"Checkpoint recovered: new P0 = {full[k][0]}, relations restored = {k}" in log
    assert sorted((tmp_path / "relations.p").read_text().split()) == sorted(" ".join(lines).split())
    res = json.loads((tmp_path / "results.json").read_text())
    assert res[0]["P0"] == item[2]
    assert res[0]["prp"] == json.loads(oracle.results_json(oit, full, "t"))["prp"]
    assert not list(tmp_path.glob("core_*.p"))


def test_cli_sigterm_keeps_checkpoint_and_resumes(tmp_path):
    """SIGTERM (Siever.cpp:67-90): no further chunk starts, core_*.p keep the finished chunks, exit
    0; the next run resumes and its relations equal an uninterrupted run's."""
    import signal
    import time
    item = "Factor=2,-25,1000000000,35359738368,3600000,5000000,2,1"  # 2^35 integers, 512 chunks
    ini = "CPU_count=0\nMemoryAllowed=64\nUseGPU=1\nGPU_sieve=1\nGPU_classes=1\n"
    full_dir, cut_dir = tmp_path / "full", tmp_path / "cut"
    full_dir.mkdir()
    cut_dir.mkdir()
    r = run_siever(full_dir, [item], ini)
    assert r.returncode == 0, r.stderr[-2000:]
    (cut_dir / "worktodo.txt").write_text(item + "\n")
    (cut_dir / "prime.ini").write_text(ini)
    proc = subprocess.Popen([BIN], cwd=cut_dir, stdin=subprocess.DEVNULL, stdout=subprocess.PIPE,
                            stderr=subprocess.PIPE, text=True)
    log = cut_dir / "siever.log"
    # the signal goes out once a chunk is checkpointed, well before the last of the 512
    t0 = time.time()
    while time.time() - t0 < 120 and proc.poll() is None and \
            not any(f.stat().st_size > 0 for f in cut_dir.glob("core_*.p")):
        time.sleep(0.002)
    proc.send_signal(signal.SIGTERM)
    out, err = proc.communicate(timeout=120)
    assert proc.returncode == 0, err[-2000:]
    text = log.read_text()
    assert "Shutdown: checkpoint preserved in core_*.p files" in text
    assert list(cut_dir.glob("core_*.p"))
    assert not (cut_dir / "results.json").exists()
    r = run_siever(cut_dir, [item], ini)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Checkpoint recovered" in (cut_dir / "siever.log").read_text()
    assert sorted((cut_dir / "relations.p").read_text().split("\n")) == \
        sorted((full_dir / "relations.p").read_text().split("\n"))
    a = json.loads((cut_dir / "results.json").read_text())
    b = json.loads((full_dir / "results.json").read_text())
    a[0].pop("Timestamp")
    b[0].pop("Timestamp")
    assert a == b
