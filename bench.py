"""Benchmark of the MI355X sieve worker: primes/sec sieved + BSGS'd at p ~ 1e12 on one GPU.

Workload (one step): the item Factor=10,111,P0=1e12,P1=1e12+2^28,N0=0,N1=1e7,mod=6,res=4 with
per-prime residue classes on, run end to end: device sieve, residue filters, BSGS, and the exact
host re-check of every flagged number, chunked as the reference chunks it with MemoryAllowed=64
and CPU_count=0 (2^26 integers per chunk).  value = primes in [P0, P1) / seconds per step.

Prints one JSON line.  The CPU baseline is the oracle (a C restatement of the reference's CPU
path) timed on a bounded slice of the same item on this machine's cores.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))

WORKLOAD = dict(b=10, c=111, P0=10**12, P1=10**12 + 2**28, N0=0, N1=10_000_000, mod=6, res=4)
CHUNK = 1 << 26
# fp64 operations per modular product of the FP64 field: mul, fma, fma + add (quotient), fma + add
FLOP_PER_PRODUCT_F64 = 9
# 32-bit multiplies per Montgomery product over two 32-bit words (per word of b: a * b_i: 2,
# m = t0 * n0: 1, m * p: 2)
MUL_PER_PRODUCT_MONT = 10
PEAK_FP64_TFLOPS = 78.6          # MI355X published FP64 vector/matrix peak
# 32x32-bit multiplies per 64-bit Montgomery product (mmul), the work unit of the integer roofline:
# 4 for the full a*b, 3 for lo*ninv mod 2^64, 4 for the full m*p
MUL_PER_MMUL = 11
# HBM bytes per tuned k_bsgs launch on this workload: FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (scale factors
# from a 1 GiB streaming copy), from profiles/k_bsgs_c2_hbm_traffic.txt
K_BSGS_TRAFFIC_BYTES = 2.896e8
TUNING_FILE = os.path.join(ROOT, "mass-prime-siever_amd", "tuning_mi355x.json")
# the autotuner's item set: other items of the same tuning keys as the timed workloads
TUNE_WORKTODO = os.path.join(ROOT, "mass-prime-siever_amd", "tune_worktodo.txt")


def load_tune():
    spec = importlib.util.spec_from_file_location("mps_tune", os.path.join(ROOT, "mass-prime-siever_amd", "tune.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def load_mps():
    spec = importlib.util.spec_from_file_location("mps", os.path.join(ROOT, "mass-prime-siever_amd", "mps.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mps"] = mod
    spec.loader.exec_module(mod)
    return mod


def prime_count(P0, P1):
    with open(os.path.join(ROOT, "tests", "golden", "prime_counts.json")) as f:
        counts = json.load(f)
    return counts[f"{P0}:{P1}"]


def host_cpu():
    """The host's CPU model, logical CPU count and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count() or 1, "affinity": affinity}


# Host threads of the CPU legs: the CPUs this process may use, at most 16 (one GPU's share of a
# GPU node's host cores).
CPU_THREADS_MAX = 16


def cpu_time_range(O, b, c, P0, P1, N0, N1, mod, res, threads):
    item = O.make_item(b, c, P0, P1, N0, N1, mod, res)
    chunk = max(1 << 20, (P1 - P0) // (4 * threads))
    t0 = time.perf_counter()
    rels, primes = O.process_range(item, P0, P1, chunk, threads)
    return time.perf_counter() - t0, primes, len(rels)


def cpu_baseline(threads, slice_ints):
    """The oracle's CPU path (segmented sieve + run_bsgs per prime, the reference's CPU branch) on
    the first slice_ints integers of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    w = WORKLOAD
    dt, primes, nrel = cpu_time_range(O, w["b"], w["c"], w["P0"], w["P0"] + slice_ints, w["N0"], w["N1"], w["mod"],
                                      w["res"], threads)
    hc = host_cpu()
    return {"value": primes / dt, "unit": "primes/s", "cores": threads, "kind": "port",
            "sample": f"first {slice_ints} integers of the workload's P range ({primes} primes, "
                      f"{nrel} relations) in {dt:.2f} s on {threads} threads",
            "cpu_model": hc["model"], "nproc": hc["nproc"], "affinity": hc["affinity"]}


def cpu_baseline_configs(threads, slice_ints=1 << 26):
    """The oracle's CPU path on c1 (whole) and the first slice_ints integers of c3 and c4, as
    primes/s; c4's sieving primes up to sqrt(P1) ~ 2e9 (sieve_small_primes, a one-off per item)
    are timed on their own first.  c5 is the serving mix (cpu_baseline_mix)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    out = {}
    for name, (b, c, P0, P1, N0, N1, mod, res, _) in CONFIGS.items():
        hi = min(P1, P0 + slice_ints)
        setup_s = None
        if name == "c4":
            t = time.perf_counter()
            O.sieve_primes(P1 - 2, P1)  # the sieving primes up to sqrt(P1), cached in the oracle
            setup_s = time.perf_counter() - t
        dt, primes, nrel = cpu_time_range(O, b, c, P0, hi, N0, N1, mod, res, threads)
        out[name] = {"value": primes / dt, "unit": "primes/s", "cores": threads, "kind": "port",
                     "sample": f"first {hi - P0} integers ({primes} primes, {nrel} relations) in {dt:.2f} s"}
        if setup_s is not None:
            out[name]["sieving_primes_setup_s"] = setup_s
    out["c5"] = cpu_baseline_mix(O, threads)
    return out


def cpu_baseline_mix(O, threads, slice_ints=1 << 20):
    """The oracle's CPU path on the serving mix (seed 0): the first slice_ints integers of each of
    the 256 items' P ranges, the items spread over the threads (the oracle's calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    items = load_mix(256)

    def one(it):
        lo = it["P0"]
        hi = min(it["P1"], lo + slice_ints)
        item = O.make_item(it["b"], it["c"], lo, hi, it["N0"], it["N1"], it["mod"], it["res"])
        rels, primes = O.process_range(item, lo, hi, 1 << 20, 1, 1 << 18)
        return primes, len(rels), hi - lo

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(one, items))
    dt = time.perf_counter() - t0
    primes, nrel, ints = (sum(r[k] for r in res) for k in range(3))
    return {"value": primes / dt, "unit": "primes/s", "cores": threads, "kind": "port",
            "sample": f"first {slice_ints} integers of each of the {len(items)} mix items ({ints} integers, "
                      f"{primes} primes, {nrel} relations) in {dt:.2f} s"}


def load_mix(n):
    with open(os.path.join(ROOT, "tests", "golden", "mix_seed0.json")) as f:
        items = json.load(f)["items"]
    return items[:n]


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(q * len(xs)))]


def drain(eng, n, log, label):
    """Finished-item records of an engine until n have finished, with a progress line every few
    seconds (the engine runs on its own threads; this thread only polls)."""
    done, t0, last = [], time.perf_counter(), time.perf_counter()
    while len(done) < n:
        time.sleep(0.002)
        done += eng.poll()
        if time.perf_counter() - last > 5:
            last = time.perf_counter()
            log(f"{label}: {len(done)}/{n} items, {last - t0:.1f} s")
    return done


def mix_items(mps, n_items):
    items = load_mix(n_items)
    primes = sum(d["primes"] for d in items)
    wis = [mps.work_item(d["b"], d["c"], d["P0"], d["P1"], d["N0"], d["N1"], d["mod"], d["res"]) for d in items]
    return items, wis, primes


def plain_mix(g, wis, log, label):
    """The plain per-chunk worker over the mix, one item after another: service times, relations."""
    svc, rels_all = [], []
    t0 = time.perf_counter()
    for k, w in enumerate(wis):
        t = time.perf_counter()
        rels, _ = g.run_item(w, chunk_ints=CHUNK)
        svc.append(time.perf_counter() - t)
        rels_all.append(rels)
        if k % 64 == 63:
            log(f"{label}: {k + 1} items, {time.perf_counter() - t0:.2f} s")
    return svc, rels_all


def plain_served(g, wis, arrivals, log):
    """The plain worker as one FIFO server under an arrival schedule, on the wall clock: an item
    starts once it has arrived and the item before it has finished; its latency runs from its
    arrival to its finish."""
    lat, rels = [], []
    start = time.perf_counter()
    for k, (a, w) in enumerate(zip(arrivals, wis)):
        delay = a - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        rels.append(g.run_item(w, chunk_ints=CHUNK)[0])
        lat.append(time.perf_counter() - start - a)
        if k % 64 == 63:
            log(f"mix plain served: {k + 1} items")
    return lat, rels


def run_mix(mps, g, n_items, load, log, plain_untuned):
    """The serving mix (seed 0): the plain per-chunk worker one item after another, the batched
    engine with every item submitted at once (throughput), then the batched engine and the plain
    worker each under the same Poisson arrivals at `load` times the plain worker's capacity
    (latency, measured from each item's arrival to its finish)."""
    items, wis, primes = mix_items(mps, n_items)
    svc, rel_plain = plain_mix(g, wis, log, "mix plain")
    t_plain = sum(svc)
    # batched, all at once
    eng = mps.Batch(g)
    t = time.perf_counter()
    ids = [eng.submit(w, tag=k) for k, w in enumerate(wis)]
    done = drain(eng, len(wis), log, "mix batched")
    t_batch = time.perf_counter() - t
    same = all(eng.relations(i) == rel_plain[k] for k, i in enumerate(ids))
    bst = eng.stats()
    eng.close()
    log(f"mix batched: {t_batch:.2f} s (plain {t_plain:.2f} s), identical={same}")
    # served: Poisson arrivals at load x plain capacity
    rate = load * len(wis) / t_plain
    arrivals, acc = [], 0.0
    for d in items:
        acc += d["gap"] / rate
        arrivals.append(acc)
    eng = mps.Batch(g)
    start = time.perf_counter()
    ids = []
    for k, w in enumerate(wis):
        delay = arrivals[k] - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        ids.append(eng.submit(w, tag=k))
        if k % 64 == 63:
            log(f"mix served: {k + 1} items submitted")
    served = drain(eng, len(wis), log, "mix served")
    served_same = all(eng.relations(i) == rel_plain[k] for k, i in enumerate(ids))
    eng.close()
    lat = [d["t_done"] - d["t_submit"] for d in served]
    log(f"mix served: p50 {pct(lat, 0.5):.3f} s p99 {pct(lat, 0.99):.3f} s, identical={served_same}")
    plain_lat, rel_plain_served = plain_served(g, wis, arrivals, log)
    served_same = served_same and rel_plain_served == rel_plain
    log(f"mix plain served: p50 {pct(plain_lat, 0.5):.3f} s p99 {pct(plain_lat, 0.99):.3f} s")
    out = {
        "workload": f"serving mix seed 0, {len(wis)} items, {primes} primes",
        "primes": primes,
        "plain_s": t_plain, "plain_primes_per_s": primes / t_plain,
        "batched_s": t_batch, "batched_primes_per_s": primes / t_batch,
        "speedup_batched_vs_plain": t_plain / t_batch,
        "relations_identical": bool(same and served_same),
        "statuses_ok": all(d["status"] == 0 for d in done + served),
        "served_load": load, "arrival_rate_items_per_s": rate,
        "served_p50_s": pct(lat, 0.5), "served_p99_s": pct(lat, 0.99),
        "plain_served_p50_s": pct(plain_lat, 0.5), "plain_served_p99_s": pct(plain_lat, 0.99),
        "batch_stats": bst,
    }
    if plain_untuned is not None:
        t_u, rels_u = plain_untuned
        out["plain_untuned_s"] = t_u
        out["speedup_batched_vs_plain_untuned"] = t_u / t_batch
        out["relations_identical"] = out["relations_identical"] and rels_u == rel_plain
    return out


CONFIGS = {  # the other configurations of the metric, as (b, c, P0, P1, N0, N1, mod, res, classes)
    "c1": (2, -25, 10**9, 10**9 + 2**26, 3_600_000, 5_000_000, 2, 1, False),
    "c3": (2, -25, 95 * 10**12, 95 * 10**12 + 2**28, 3_600_000, 5_000_000, 2, 1, True),
    "c4": (3, -2, 4 * 10**18, 4 * 10**18 + 2**30, 0, 10_000_000, 1, 0, True),
}


def item_stats(st, reps):
    """Per-item device counts and stage times of `reps` runs' accumulated stats."""
    return {"entries": st["entries"] // reps, "candidates": st["candidates"] // reps,
            "survivors": st["survivors"] // reps, "flagged": st["flagged"] // reps,
            "products": {k: st[k] // reps for k in ("w_fermat", "w_setup", "w_bsgs")},
            "stage_ms": {k: st[k] / reps for k in ("ms_sieve", "ms_filter", "ms_setup", "ms_bsgs")},
            "field": "fp64" if st["field"] == 1 else "montgomery"}


def int_roofline(st, reps, t_plain, t_batch, rates):
    """The item's modular products W (Fermat, setup, baby and giant steps, counted on the device)
    as 32x32-bit multiplies (MUL_PER_MMUL each) per second of item time, against the measured
    v_mad_u64_u32 issue rate of the chip."""
    W = (st["w_fermat"] + st["w_setup"] + st["w_bsgs"]) / reps
    peak = rates["mad_u64_u32_per_s"]
    return {"W_mmul": W, "unit": "32x32 multiplies/s", "peak": peak,
            "achieved_plain": MUL_PER_MMUL * W / t_plain, "frac_plain": MUL_PER_MMUL * W / t_plain / peak,
            "achieved_batched": MUL_PER_MMUL * W / t_batch, "frac_batched": MUL_PER_MMUL * W / t_batch / peak}


def kernel_roofline(st, rates):
    """k_bsgs alone, from its event times: FP64 field products x 9 FLOP against the FP64 peak, or
    Montgomery products x 10 32-bit multiplies against the measured v_mad_u64_u32 rate."""
    launches = max(1, st["n_bsgs_launches"])
    ms = st["ms_bsgs"] / launches
    prod = st["w_bsgs"] / launches
    if st["field"] == 1:
        achieved = prod * FLOP_PER_PRODUCT_F64 / (ms * 1e-3)
        return {"kernel": "k_bsgs<F64Field>", "bound": "valu", "unit": "FLOP/s", "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS * 1e12, "frac": achieved / (PEAK_FP64_TFLOPS * 1e12),
                "ms_per_launch": ms, "products_per_launch": prod}
    achieved = prod * MUL_PER_PRODUCT_MONT / (ms * 1e-3)
    return {"kernel": "k_bsgs<MontField>", "bound": "valu", "unit": "32x32 multiplies/s", "achieved": achieved,
            "peak": rates["mad_u64_u32_per_s"], "frac": achieved / rates["mad_u64_u32_per_s"],
            "ms_per_launch": ms, "products_per_launch": prod}


def run_configs(mps, g, log, rates, reps=3):
    """Item latency and primes/s of c1, c3 and c4 on the plain worker and the batched engine."""
    out = {}
    for name, (b, c, P0, P1, N0, N1, mod, res, classes) in CONFIGS.items():
        w = mps.work_item(b, c, P0, P1, N0, N1, mod, res)
        g.set_classes(classes)
        try:
            g.run_item(w, chunk_ints=CHUNK)  # warm-up
            t_plain = []
            g.reset_stats()
            for _ in range(reps):
                t = time.perf_counter()
                rels, _ = g.run_item(w, chunk_ints=CHUNK)
                t_plain.append(time.perf_counter() - t)
            st = g.stats()
            # the kernels alone on the device (one feeder: passes one after another) for k_bsgs's
            # own launch time; with two feeders the passes of two chunks overlap
            g.reset_stats()
            rels_serial, _ = g.run_item(w, chunk_ints=CHUNK, feeder_threads=1)
            st_serial = g.stats()
            eng = mps.Batch(g)
            t_batch, same = [], True
            for _ in range(reps):  # the first builds the slots' graphs
                t = time.perf_counter()
                i = eng.submit(w)
                eng.wait_all()
                t_batch.append(time.perf_counter() - t)
                done = eng.poll()
                same = same and done[0]["status"] == 0 and eng.relations(i) == rels
            eng.close()
        finally:
            g.set_classes(True)
        n = prime_count(P0, P1)
        out[name] = {"item": f"Factor={b},{c},{P0},{P1},{N0},{N1}" + (f",{mod},{res}" if mod > 1 else "") +
                             f" GPU_classes={int(classes)}",
                     "primes": n, "relations": len(rels),
                     "plain_item_s": min(t_plain), "plain_primes_per_s": n / min(t_plain),
                     "batched_item_s": min(t_batch), "batched_primes_per_s": n / min(t_batch),
                     "relations_identical": same and rels_serial == rels,
                     "device": item_stats(st, reps),
                     "int_roofline": int_roofline(st, reps, min(t_plain), min(t_batch), rates),
                     "kernel_roofline": kernel_roofline(st_serial, rates)}
        log(f"{name}: plain {min(t_plain)*1e3:.1f} ms, batched {min(t_batch)*1e3:.1f} ms per item")
    return out


def time_plain(g, item, steps, warmup, feeders=0):
    """Wall time of `steps` runs of the item on the plain worker, after `warmup` runs; feeders = the
    worker's feeder threads (0: the library's default, 2)."""
    for _ in range(warmup):
        g.run_item(item, chunk_ints=CHUNK, feeder_threads=feeders)
    g.reset_stats()
    times, rels = [], None
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        rels, _ = g.run_item(item, chunk_ints=CHUNK, feeder_threads=feeders)
        times.append(time.perf_counter() - t0)
    ms = 1000.0 * (time.perf_counter() - t_all) / steps
    return ms, times, rels, g.stats()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--feeders", type=int, default=0,
                    help="feeder threads of the plain worker's timed legs (0: the library's default, 2; "
                         "1 runs the passes one after another, for kernel profiles)")
    ap.add_argument("--cpu-slice", type=int, default=1 << 27)
    ap.add_argument("--mix", type=int, default=256, help="items of the serving mix (0 = skip the mix)")
    ap.add_argument("--mix-load", type=float, default=0.7, help="arrival rate / plain worker capacity")
    ap.add_argument("--tuning", default=TUNING_FILE,
                    help="tuning table to load after the untuned legs ('none': keep the defaults)")
    ap.add_argument("--tune", action="store_true", help="autotune the workloads' keys live instead of loading")
    ap.add_argument("--no-configs", action="store_true", help="skip the c1/c3/c4 item legs")
    ap.add_argument("--plain-only", action="store_true",
                    help="only the tuned plain c2 leg (no untuned or batched legs): for kernel profiles")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB streaming copy (k_copy), the byte-count reference of HBM counters")
    ap.add_argument("--profile-config", choices=sorted(CONFIGS), default=None,
                    help="only `steps` tuned plain runs of this configuration (for kernel profiles)")
    ap.add_argument("--no-cpu-configs", action="store_true", help="skip the CPU legs of c1, c3 and c4")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this worker runs on one GPU; --gpus must be 1 "
                                   "(P ranges shard across GPUs as independent processes)"}))
        sys.exit(2)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    mps = load_mps()
    tune = load_tune()
    w = WORKLOAD
    g = mps.GpuSiever()
    if not g.init(0):
        raise SystemExit(g.banner)
    g.clear_tuning()
    if args.profile_config:
        tune.load(g, args.tuning)
        b, c, P0, P1, N0, N1, mod, res, classes = CONFIGS[args.profile_config]
        g.set_classes(classes)
        it = mps.work_item(b, c, P0, P1, N0, N1, mod, res)
        ms, _, rels, st = time_plain(g, it, args.steps, args.warmup, args.feeders)
        log(f"{args.profile_config} plain {ms:.2f} ms/item")
        print(json.dumps({"config": args.profile_config, "ms_per_item": ms, "relations": len(rels),
                          "device": item_stats(st, args.steps)}), flush=True)
        return
    rates = mps.measure_valu(0)
    log(f"VALU rates: mad_u64_u32 {rates['mad_u64_u32_per_s']:.3e}/s, fma64 {rates['fma64_per_s']:.3e}/s")
    item = mps.work_item(w["b"], w["c"], w["P0"], w["P1"], w["N0"], w["N1"], w["mod"], w["res"])
    nprimes = prime_count(w["P0"], w["P1"])
    copy_gbps = mps.copy_probe(1 << 30, 3) if args.copy_probe else None
    tuned = args.tune or args.tuning != "none"

    # untuned legs first (the handle's default shapes)
    ms_untuned, rels_untuned, mix_untuned = None, None, None
    if tuned and not args.plain_only:
        ms_untuned, _, rels_untuned, _ = time_plain(g, item, args.steps, args.warmup)
        log(f"c2 plain untuned {ms_untuned:.2f} ms/step")
        if args.mix:
            _, wis_u, _ = mix_items(mps, args.mix)
            svc_u, rels_mix_u = plain_mix(g, wis_u, log, "mix plain untuned")
            mix_untuned = (sum(svc_u), rels_mix_u)
    tuning_info = {"source": "defaults"}
    if args.tune:
        # tuned live on the held-out set (none of its items is timed here)
        with open(TUNE_WORKTODO) as f:
            items = tune.parse_worktodo(f.read())
        t = time.perf_counter()
        table = tune.autotune_items(g, items, log=log)
        tuning_info = {"source": "autotuned live on " + os.path.relpath(TUNE_WORKTODO, ROOT), "keys": len(table),
                       "tune_s": time.perf_counter() - t}
        table_keys = {r["key"] for r in table}
    elif args.tuning != "none":
        if not os.path.exists(args.tuning):  # a missing table would leave the tuned legs untuned
            raise SystemExit(f"tuning table {args.tuning} not found")
        tuning_info = {"source": os.path.relpath(args.tuning, ROOT), "keys": tune.load(g, args.tuning)}
        with open(args.tuning) as f:
            table_keys = {r["key"] for r in json.load(f)["entries"]}
    if tuned:
        # how much of the timed work has a tuned key (the rest runs with the defaults)
        tuning_info["c2_key_tuned"] = g.get_tuning(item)[1] in table_keys
        if args.mix:
            _, wis_t, _ = mix_items(mps, args.mix)
            tuning_info["mix_items_with_tuned_key"] = sum(g.get_tuning(x)[1] in table_keys for x in wis_t)
    log(f"tuning: {tuning_info}")

    ms_per_step, times, rels, s = time_plain(g, item, args.steps, args.warmup, args.feeders)
    log(f"c2 plain {ms_per_step:.2f} ms/step")
    # the same item with one feeder: its passes run one after another, each kernel alone on the
    # device, so the HIP-event times are the kernels' own (with two feeders two chunks' passes
    # overlap, and a launch's event time includes the other pass's kernels)
    if args.feeders == 1:
        ms_serial, rels_serial, s_serial = ms_per_step, rels, s
    else:
        ms_serial, _, rels_serial, s_serial = time_plain(g, item, args.steps, 1, feeders=1)
        log(f"c2 plain, one feeder {ms_serial:.2f} ms/step")

    # roofline of the dominant kernel (k_bsgs), from HIP events on the worker's stream
    kr = kernel_roofline(s_serial, rates)
    roof = {"bound": "valu", "achieved": kr["achieved"] / 1e12, "peak": kr["peak"] / 1e12,
            "unit": "TFLOP/s" if s["field"] == 1 else "T 32x32 multiplies/s", "frac": kr["frac"],
            "traffic": K_BSGS_TRAFFIC_BYTES, "kernel": kr["kernel"], "ms_per_launch": kr["ms_per_launch"],
            "products_per_launch": kr["products_per_launch"], "tuning": g.get_tuning(item)[0]}

    line = {
        "metric": "primes/sec sieved+BSGS'd at p~1e12 (1 GPU)",
        "value": nprimes / (ms_per_step / 1000.0),
        "unit": "primes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if s["field"] == 0 else "f64-exact",
        "data": "synthetic (deterministic work item)",
        "config": {"workload": "Factor=10,111,1000000000000,1000268435456,0,10000000,6,4 GPU_classes=1",
                   "chunk_ints": CHUNK, "primes": nprimes, "relations": len(rels),
                   "tuning": tuning_info,
                   "p50_step_ms": sorted(times)[len(times) // 2] * 1000,
                   "feeders": args.feeders or 2,
                   "serial_ms_per_step": ms_serial,
                   "serial_relations_identical": rels_serial == rels,
                   # per-stage HIP-event times of the one-feeder leg (stages in sequence) and of the
                   # timed leg (two chunks' passes overlap, so its stages add up to more than a step)
                   "stage_ms_per_step": {k: s_serial[k] / args.steps for k in ("ms_sieve", "ms_filter", "ms_setup",
                                                                                 "ms_bsgs")},
                   "stage_ms_overlapped_per_step": {k: s[k] / args.steps for k in ("ms_sieve", "ms_filter",
                                                                                     "ms_setup", "ms_bsgs")},
                   "entries_per_step": s["entries"] // args.steps,
                   "survivors_per_step": s["survivors"] // args.steps,
                   "candidates_per_step": s["candidates"] // args.steps,
                   "products_per_step": {k: s[k] // args.steps for k in ("w_fermat", "w_setup", "w_bsgs")},
                   "flagged_per_step": s["flagged"] // args.steps},
        "roofline": roof,
        "int_roofline": int_roofline(s, args.steps, ms_per_step / 1000.0, ms_per_step / 1000.0, rates),
        "valu_rates": rates,
    }
    if ms_untuned is not None:
        line["config"]["untuned_ms_per_step"] = ms_untuned
        line["config"]["untuned_relations_identical"] = rels_untuned == rels

    if args.plain_only:
        print(json.dumps(line), flush=True)
        return
    # the same item through the batched engine (its passes spread over the engine's slots)
    # (wait_all returns when the engine's last item finishes: no polling interval in the timing)
    eng = mps.Batch(g)
    for _ in range(max(1, args.warmup)):
        eng.submit(item)
    eng.wait_all()
    eng.poll()
    t = time.perf_counter()
    bids = [eng.submit(item) for _ in range(args.steps)]
    eng.wait_all()
    ms_batched = 1000.0 * (time.perf_counter() - t) / args.steps
    eng.poll()
    same = all(eng.relations(i) == rels for i in bids)
    eng.close()
    line["config"]["batched_ms_per_step"] = ms_batched
    line["config"]["batched_primes_per_s"] = nprimes / (ms_batched / 1000.0)
    line["config"]["batched_relations_identical"] = same
    log(f"c2 plain {ms_per_step:.2f} ms/step, batched {ms_batched:.2f} ms/step")
    if copy_gbps is not None:
        line["config"]["hbm_copy_gbps"] = copy_gbps
    if not args.no_configs:
        line["configs"] = run_configs(mps, g, log, rates)
    if args.mix:
        line["mix"] = run_mix(mps, g, args.mix, args.mix_load, log, mix_untuned)
        line["p50_item_latency_s"] = line["mix"]["served_p50_s"]
        line["p99_item_latency_s"] = line["mix"]["served_p99_s"]
    if not args.no_cpu:
        threads = max(1, min(CPU_THREADS_MAX, host_cpu()["affinity"]))
        line["cpu_baseline"] = cpu_baseline(threads, args.cpu_slice)
        log(f"cpu baseline: {line['cpu_baseline']['value']:.3e} primes/s on {threads} threads")
        if not args.no_cpu_configs:
            line["cpu_baseline_configs"] = cpu_baseline_configs(threads)
            log(f"cpu baseline configs: { {k: round(v['value']) for k, v in line['cpu_baseline_configs'].items()} }")
    print(json.dumps(line), flush=True)
    g.close()


if __name__ == "__main__":
    main()
