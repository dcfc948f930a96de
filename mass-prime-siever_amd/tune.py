"""Autotuner driver: picks the BSGS shape (wavefronts per entry, baby-step caps, LDS table load)
for every kind of item in a worktodo.txt file and writes a tuning table that a GpuSiever loads.

    python tune.py worktodo.txt --out tuning.json [--sample-odds 16777216]

Kinds are the library's tuning keys (class mode and class-modulus kind, p magnitude, exponent
span); one item per key is timed (include/mps_gpu.h mps_gpu_autotune).  The table lists, per key,
a representative item and the shape, so load() can install it on any handle.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def _mps():
    if "mps" in sys.modules:
        return sys.modules["mps"]
    import importlib.util
    spec = importlib.util.spec_from_file_location("mps", os.path.join(HERE, "mps.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mps"] = mod
    spec.loader.exec_module(mod)
    return mod


def parse_worktodo(text):
    """Factor=b,c,P0,P1,N0,N1[,mod,res] lines (the reference's worktodo.txt) as tuples."""
    out = []
    for line in text.splitlines():
        line = line.strip()
        if not line.startswith("Factor="):
            continue
        f = [int(x) for x in line[len("Factor="):].split(",")]
        b, c, P0, P1, N0, N1 = f[:6]
        mod, res = (f[6], f[7]) if len(f) >= 8 else (1, 0)
        out.append((b, c, P0, P1, N0, N1, mod, res))
    return out


def autotune_items(g, items, sample_odds=1 << 24, log=None):
    """Tunes one representative item per key; returns the table records."""
    mps = _mps()
    reps = {}
    for it in items:
        w = mps.work_item(*it)
        reps.setdefault(g.get_tuning(w)[1], it)
    table = []
    for key, it in reps.items():
        t0 = time.perf_counter()
        best, best_ms, default_ms = g.autotune(mps.work_item(*it), sample_odds)
        rec = {"key": key, "item": list(it), "tuning": best, "best_ms": best_ms, "default_ms": default_ms,
               "tune_s": time.perf_counter() - t0}
        table.append(rec)
        if log:
            log(f"key {key:#x} item {it}: {best} {best_ms:.3f} ms (default {default_ms:.3f} ms)")
    return table


def save(path, table, device=""):
    with open(path, "w") as f:
        json.dump({"device": device, "entries": table}, f, indent=1)
        f.write("\n")


def load(g, path):
    """Installs every entry of a tuning table on handle g; returns the number installed."""
    mps = _mps()
    with open(path) as f:
        table = json.load(f)["entries"]
    for rec in table:
        g.set_tuning(mps.work_item(*rec["item"]), **rec["tuning"])
    return len(table)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("worktodo")
    ap.add_argument("--out", required=True)
    ap.add_argument("--sample-odds", type=int, default=1 << 24)
    args = ap.parse_args()
    mps = _mps()
    g = mps.GpuSiever()
    if not g.init(0):
        raise SystemExit(g.banner)
    items = parse_worktodo(open(args.worktodo).read())
    t0 = time.perf_counter()
    table = autotune_items(g, items, args.sample_odds, log=lambda m: print(m, flush=True))
    save(args.out, table, device=g.banner)
    print(f"{len(table)} keys from {len(items)} items in {time.perf_counter() - t0:.1f} s -> {args.out}", flush=True)
    g.close()


if __name__ == "__main__":
    main()
