"""Python view of libmps.so, the MI355X GPU worker (include/mps_gpu.h, include/mps_host.h).

GpuSiever mirrors the reference's chunk interface (Gpu.hpp:1040-1074): init/ok/filter/
set_classes/sieve_filter with the same argument meaning; a declined chunk returns False
(Gpu.hpp:857-979) and a device error disables the GPU sieve for good.  run_item is the GPU
branch of the reference's chunk loop plus the exact host re-check.  There is no CPU fallback:
loading fails loudly when libmps.so is missing.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmps.so")

MPS_OK, MPS_DECLINED, MPS_ERROR, MPS_CAPACITY, MPS_STOPPED = 0, 1, 2, 3, 4


class Item(ctypes.Structure):
    """mps_item: the fields of a WorkItem the device needs (Types.hpp:14-26)."""
    _fields_ = [("b", ctypes.c_int64), ("c", ctypes.c_int64), ("N0", ctypes.c_int64), ("N1", ctypes.c_int64),
                ("cls_mod", ctypes.c_uint64), ("cls_res", ctypes.c_uint64)]


class WorkItem(ctypes.Structure):
    """mps_work_item (Types.hpp:14-26)."""
    _fields_ = [("b", ctypes.c_int64), ("c", ctypes.c_int64), ("P0", ctypes.c_uint64), ("P1", ctypes.c_uint64),
                ("N0", ctypes.c_int64), ("N1", ctypes.c_int64), ("cls_mod", ctypes.c_uint64),
                ("cls_res", ctypes.c_uint64), ("P0_original", ctypes.c_uint64)]


class Relation(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint64), ("M", ctypes.c_uint64), ("q", ctypes.c_uint64)]


class DevRelation(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint64), ("M", ctypes.c_uint64)]


HOST_DECIDES = 0xFFFFFFFFFFFFFFFF
NOALLOC_MAX_INTS = 1 << 20  # MPS_NOALLOC_MAX_INTS (include/mps_gpu.h)


class RunOpts(ctypes.Structure):
    _fields_ = [("chunk_ints", ctypes.c_uint64), ("feeder_threads", ctypes.c_int), ("on_chunk", ctypes.c_void_p),
                ("ctx", ctypes.c_void_p), ("stop", ctypes.c_void_p)]


class RunStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("device_call_s", ctypes.c_double), ("check_s", ctypes.c_double),
                ("chunks", ctypes.c_uint64), ("declined", ctypes.c_uint64), ("flagged", ctypes.c_uint64),
                ("prime_flagged", ctypes.c_uint64), ("relations", ctypes.c_uint64)]


class GpuStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("passes", "odds", "survivors", "candidates", "entries", "flagged",
                                               "flag2", "fermat_fail", "w_fermat", "w_setup", "w_bsgs")] + \
               [(n, ctypes.c_double) for n in ("ms_sieve", "ms_filter", "ms_setup", "ms_bsgs")] + \
               [("n_bsgs_launches", ctypes.c_uint64), ("field", ctypes.c_int), ("tbits", ctypes.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Tuning(ctypes.Structure):
    """mps_tuning (include/mps_gpu.h): the BSGS shape of one kind of item."""
    _fields_ = [("waves", ctypes.c_uint32), ("mcap", ctypes.c_uint32), ("cmax", ctypes.c_uint32),
                ("load", ctypes.c_uint32), ("policy", ctypes.c_uint32), ("ext", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class BatchOpts(ctypes.Structure):
    """mps_batch_opts (include/mps_batch.h)."""
    _fields_ = [("slots", ctypes.c_int), ("pass_odds", ctypes.c_uint64), ("use_graphs", ctypes.c_int),
                ("policy", ctypes.c_int), ("max_active", ctypes.c_int), ("mem_budget", ctypes.c_uint64)]


class ItemDone(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int),
                ("relations", ctypes.c_uint64), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("primes_flagged", ctypes.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("passes", "items_submitted", "items_done", "graph_launches",
                                               "graphs_built")] + \
               [(n, ctypes.c_double) for n in ("busy_s", "device_s", "host_s")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ValuRates(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("fma64_per_s", "mul_lo32_per_s", "mul_hi32_per_s",
                                               "mad_u64_u32_per_s", "add32_per_s", "fma64_latency_cycles",
                                               "clock_mhz")] + [("cus", ctypes.c_int)]


def measure_valu(device_index=0):
    """Measured VALU issue rates of the device (include/mps_gpu.h mps_gpu_measure_valu)."""
    r = ValuRates()
    if lib().mps_gpu_measure_valu(device_index, ctypes.byref(r)) != MPS_OK:
        raise RuntimeError("VALU rate probe failed")
    return {n: getattr(r, n) for n, _ in r._fields_}


def copy_probe(bytes_=1 << 30, reps=5, device_index=0):
    """Achieved HBM copy rate (read + write, GB/s) of a plain streaming kernel."""
    gbps = ctypes.c_double()
    if lib().mps_gpu_copy_probe(device_index, bytes_, reps, ctypes.byref(gbps)) != MPS_OK:
        raise RuntimeError("copy probe failed")
    return gbps.value


def build():
    """Compiles libmps.so for gfx950 (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


EXPORTS = ["mps_gpu_create", "mps_gpu_ok", "mps_gpu_set_classes", "mps_gpu_set_class_policy", "mps_gpu_sieve_filter", "mps_gpu_filter",
           "mps_gpu_sieve_relations", "mps_gpu_sieve_relations_noalloc",
           "mps_gpu_last_error", "mps_gpu_destroy", "mps_gpu_get_stats", "mps_gpu_reset_stats",
           "mps_gpu_configure", "mps_gpu_debug_stage", "mps_run_item", "mps_results_json", "mps_is_prime64",
           "mps_multiplicative_order", "mps_run_bsgs", "mps_gpu_measure_valu", "mps_gpu_copy_probe", "mps_gpu_autotune", "mps_gpu_get_tuning",
           "mps_gpu_set_tuning", "mps_gpu_clear_tuning", "mps_batch_create",
           "mps_batch_submit", "mps_batch_poll", "mps_batch_wait_all", "mps_batch_relations", "mps_batch_clock",
           "mps_batch_get_stats", "mps_batch_last_error", "mps_batch_destroy"]

_lib = None


def lib():
    """Loads libmps.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (mps.build() or make -C {HERE})")
    L = ctypes.CDLL(LIB_PATH)
    u64, i32, sz, vp = ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p
    P = ctypes.POINTER
    L.mps_gpu_create.restype = vp
    L.mps_gpu_create.argtypes = [i32, ctypes.c_char_p, sz]
    L.mps_gpu_ok.restype = i32
    L.mps_gpu_ok.argtypes = [vp]
    L.mps_gpu_set_classes.restype = None
    L.mps_gpu_set_classes.argtypes = [vp, i32]
    L.mps_gpu_set_class_policy.restype = i32
    L.mps_gpu_set_class_policy.argtypes = [vp, i32]
    L.mps_gpu_sieve_filter.restype = i32
    L.mps_gpu_sieve_filter.argtypes = [vp, u64, u64, P(Item), vp, sz, P(sz)]
    L.mps_gpu_sieve_relations.restype = i32
    L.mps_gpu_sieve_relations.argtypes = [vp, u64, u64, P(Item), vp, sz, P(sz)]
    L.mps_gpu_sieve_relations_noalloc.restype = i32
    L.mps_gpu_sieve_relations_noalloc.argtypes = [vp, u64, u64, P(Item), vp, sz, P(sz)]
    L.mps_gpu_filter.restype = i32
    L.mps_gpu_filter.argtypes = [vp, vp, sz, P(Item), vp, sz, P(sz)]
    L.mps_gpu_last_error.restype = ctypes.c_char_p
    L.mps_gpu_last_error.argtypes = [vp]
    L.mps_gpu_destroy.restype = None
    L.mps_gpu_destroy.argtypes = [vp]
    L.mps_gpu_get_stats.restype = i32
    L.mps_gpu_get_stats.argtypes = [vp, P(GpuStats)]
    L.mps_gpu_reset_stats.restype = None
    L.mps_gpu_reset_stats.argtypes = [vp]
    L.mps_gpu_configure.restype = i32
    L.mps_gpu_configure.argtypes = [vp, i32, u64, ctypes.c_uint32, ctypes.c_uint32]
    L.mps_gpu_debug_stage.restype = i32
    L.mps_gpu_debug_stage.argtypes = [vp, i32, u64, u64, P(Item), vp, sz, P(sz)]
    L.mps_run_item.restype = i32
    L.mps_run_item.argtypes = [vp, P(WorkItem), P(RunOpts), vp, sz, P(sz), P(RunStats)]
    L.mps_results_json.restype = sz
    L.mps_results_json.argtypes = [P(WorkItem), vp, sz, ctypes.c_char_p, ctypes.c_char_p, sz]
    L.mps_is_prime64.restype = i32
    L.mps_is_prime64.argtypes = [u64]
    L.mps_multiplicative_order.restype = u64
    L.mps_multiplicative_order.argtypes = [u64, u64]
    L.mps_run_bsgs.restype = i32
    L.mps_run_bsgs.argtypes = [u64, P(WorkItem), P(Relation)]
    L.mps_gpu_measure_valu.restype = i32
    L.mps_gpu_measure_valu.argtypes = [i32, P(ValuRates)]
    L.mps_gpu_autotune.restype = i32
    L.mps_gpu_autotune.argtypes = [vp, P(Item), u64, u64, P(Tuning), P(ctypes.c_double), P(ctypes.c_double)]
    L.mps_gpu_get_tuning.restype = i32
    L.mps_gpu_get_tuning.argtypes = [vp, P(Item), u64, P(Tuning), P(ctypes.c_uint32)]
    L.mps_gpu_set_tuning.restype = i32
    L.mps_gpu_set_tuning.argtypes = [vp, P(Item), u64, P(Tuning)]
    L.mps_gpu_clear_tuning.restype = None
    L.mps_gpu_clear_tuning.argtypes = [vp]
    L.mps_gpu_copy_probe.restype = i32
    L.mps_gpu_copy_probe.argtypes = [i32, u64, i32, P(ctypes.c_double)]
    L.mps_batch_create.restype = vp
    L.mps_batch_create.argtypes = [vp, P(BatchOpts)]
    L.mps_batch_submit.restype = ctypes.c_int64
    L.mps_batch_submit.argtypes = [vp, P(WorkItem), u64]
    L.mps_batch_poll.restype = i32
    L.mps_batch_poll.argtypes = [vp, P(ItemDone), sz, P(sz)]
    L.mps_batch_wait_all.restype = i32
    L.mps_batch_wait_all.argtypes = [vp]
    L.mps_batch_relations.restype = i32
    L.mps_batch_relations.argtypes = [vp, u64, vp, sz, P(sz)]
    L.mps_batch_clock.restype = ctypes.c_double
    L.mps_batch_clock.argtypes = [vp]
    L.mps_batch_get_stats.restype = i32
    L.mps_batch_get_stats.argtypes = [vp, P(BatchStats)]
    L.mps_batch_last_error.restype = ctypes.c_char_p
    L.mps_batch_last_error.argtypes = [vp]
    L.mps_batch_destroy.restype = None
    L.mps_batch_destroy.argtypes = [vp]
    _lib = L
    return L


def work_item(b, c, P0, P1, N0, N1, mod=1, res=0):
    return WorkItem(b, c, P0, P1, N0, N1, mod, res, P0)


def device_item(w):
    return Item(w.b, w.c, w.N0, w.N1, w.cls_mod, w.cls_res)


class GpuSiever:
    """The reference's GpuSiever (Gpu.hpp:1040-1074) over the C-ABI."""

    def __init__(self):
        self.h = None
        self.banner = ""

    def init(self, device_index=0, backend="auto"):
        if backend not in ("auto", "hip"):
            self.banner = f"GPU disabled: GPU_backend={backend} is not available in this build"
            return False
        buf = ctypes.create_string_buffer(512)
        h = lib().mps_gpu_create(device_index, buf, 512)
        self.banner = buf.value.decode()
        if not h:
            return False
        self.h = h
        return True

    def ok(self):
        return bool(self.h) and bool(lib().mps_gpu_ok(self.h))

    def set_classes(self, on):
        lib().mps_gpu_set_classes(self.h, 1 if on else 0)

    def set_class_policy(self, policy):
        """0: classes for an even class modulus only (the reference's rule); 1: for any (default)."""
        if lib().mps_gpu_set_class_policy(self.h, policy) != MPS_OK:
            raise ValueError(f"class policy {policy}")

    def configure(self, field=-1, pass_odds=0, mcap=0, waves=0):
        return lib().mps_gpu_configure(self.h, field, pass_odds, mcap, waves)

    def sieve_filter(self, lo, hi, item, flagged):
        """Appends what the device could not rule out in [lo|1, hi) to `flagged`; False = declined."""
        di = device_item(item) if isinstance(item, WorkItem) else item
        cap = 1 << 16
        while True:
            out = np.empty(cap, dtype=np.uint64)
            n = ctypes.c_size_t()
            st = lib().mps_gpu_sieve_filter(self.h, lo, hi, ctypes.byref(di), out.ctypes.data, cap, ctypes.byref(n))
            if st == MPS_CAPACITY:
                cap = n.value + 1024
                continue
            if st != MPS_OK:
                return False
            flagged.extend(int(x) for x in out[: n.value])
            return True

    def sieve_relations(self, lo, hi, item):
        """Exact device records (p, M) of [lo|1, hi); M == HOST_DECIDES: the host decides p.
        Returns None when the chunk is declined."""
        di = device_item(item) if isinstance(item, WorkItem) else item
        cap = 1 << 14
        while True:
            out = (DevRelation * cap)()
            n = ctypes.c_size_t()
            st = lib().mps_gpu_sieve_relations(self.h, lo, hi, ctypes.byref(di), out, cap, ctypes.byref(n))
            if st == MPS_CAPACITY:
                cap = n.value + 1024
                continue
            if st != MPS_OK:
                return None
            return [(out[i].p, out[i].M) for i in range(n.value)]

    def sieve_relations_noalloc(self, lo, hi, item, cap=1 << 14):
        """mps_gpu_sieve_relations_noalloc: (status, records) for a range of at most
        NOALLOC_MAX_INTS integers, computed on the library's per-GPU scratch array."""
        di = device_item(item) if isinstance(item, WorkItem) else item
        out = (DevRelation * cap)()
        n = ctypes.c_size_t()
        st = lib().mps_gpu_sieve_relations_noalloc(self.h, lo, hi, ctypes.byref(di), out, cap, ctypes.byref(n))
        return st, [(out[i].p, out[i].M) for i in range(min(n.value, cap))], n.value

    def filter(self, primes, item, todo):
        """Appends the ascending indices of the primes that may divide an in-class candidate."""
        di = device_item(item) if isinstance(item, WorkItem) else item
        arr = np.ascontiguousarray(np.asarray(primes, dtype=np.uint64))
        cap = max(1024, len(arr) // 16)
        while True:
            out = np.empty(cap, dtype=np.uint64)
            n = ctypes.c_size_t()
            st = lib().mps_gpu_filter(self.h, arr.ctypes.data, len(arr), ctypes.byref(di), out.ctypes.data, cap,
                                      ctypes.byref(n))
            if st == MPS_CAPACITY:
                cap = n.value + 1024
                continue
            if st != MPS_OK:
                return False
            todo.extend(int(x) for x in out[: n.value])
            return True

    def debug_stage(self, stage, lo, hi, item):
        di = device_item(item) if isinstance(item, WorkItem) else item
        width = (1, 2, 4)[stage]
        cap = ((hi - lo) // 2 + 16) * width
        out = np.empty(cap, dtype=np.uint64)
        n = ctypes.c_size_t()
        st = lib().mps_gpu_debug_stage(self.h, stage, lo, hi, ctypes.byref(di), out.ctypes.data, cap,
                                       ctypes.byref(n))
        if st != MPS_OK:
            raise RuntimeError(f"debug_stage {stage}: status {st}: {self.last_error()}")
        return out[: n.value * width].reshape(-1, width) if width > 1 else out[: n.value]

    def run_item(self, item, chunk_ints=0, feeder_threads=0):
        """Relations (sorted by p) of one item, plus the run statistics."""
        cap = 1 << 16
        while True:
            rels = (Relation * cap)()
            n = ctypes.c_size_t()
            st = RunStats()
            opts = RunOpts(chunk_ints, feeder_threads, None, None, None)
            rc = lib().mps_run_item(self.h, ctypes.byref(item), ctypes.byref(opts), rels, cap, ctypes.byref(n),
                                    ctypes.byref(st))
            if rc == MPS_CAPACITY:
                cap = n.value + 1024
                continue
            if rc != MPS_OK:
                raise RuntimeError(f"run_item status {rc}: {self.last_error()}")
            return [(rels[i].p, rels[i].M, rels[i].q) for i in range(n.value)], st

    def autotune(self, item, sample_odds=1 << 24):
        """Times the BSGS shapes on a sample pass of a work item and installs the fastest for its
        key; returns (tuning dict, best ms, default ms)."""
        t, best, dflt = Tuning(), ctypes.c_double(), ctypes.c_double()
        rc = lib().mps_gpu_autotune(self.h, ctypes.byref(device_item(item)), item.P0, sample_odds, ctypes.byref(t),
                                    ctypes.byref(best), ctypes.byref(dflt))
        if rc != MPS_OK:
            raise RuntimeError(f"autotune status {rc}: {self.last_error()}")
        return t.as_dict(), best.value, dflt.value

    def get_tuning(self, item):
        t, key = Tuning(), ctypes.c_uint32()
        lib().mps_gpu_get_tuning(self.h, ctypes.byref(device_item(item)), item.P0, ctypes.byref(t), ctypes.byref(key))
        return t.as_dict(), key.value

    def set_tuning(self, item, **kw):
        t = Tuning(*(kw.get(n, 0) for n, _ in Tuning._fields_))
        if lib().mps_gpu_set_tuning(self.h, ctypes.byref(device_item(item)), item.P0, ctypes.byref(t)) != MPS_OK:
            raise ValueError(f"tuning {kw}")

    def clear_tuning(self):
        lib().mps_gpu_clear_tuning(self.h)

    def stats(self):
        s = GpuStats()
        lib().mps_gpu_get_stats(self.h, ctypes.byref(s))
        return s.as_dict()

    def reset_stats(self):
        lib().mps_gpu_reset_stats(self.h)

    def last_error(self):
        return lib().mps_gpu_last_error(self.h).decode() if self.h else "no handle"

    def close(self):
        if self.h:
            lib().mps_gpu_destroy(self.h)
            self.h = None


class Batch:
    """The batched / served engine (include/mps_batch.h) on a GpuSiever's device.

    submit() queues an item and returns at once; poll() returns the items finished since the last
    call; relations(id) the sorted (p, M, q) of a finished item.  Items admitted together share
    the device's slots (continuous batching)."""

    def __init__(self, siever, slots=0, pass_odds=0, use_graphs=True, policy=0, max_active=0, mem_budget=0):
        self.opts = BatchOpts(slots, pass_odds, 1 if use_graphs else 0, policy, max_active, mem_budget)
        self.h = lib().mps_batch_create(siever.h, ctypes.byref(self.opts))
        if not self.h:
            raise RuntimeError(f"batch engine: {siever.last_error()}")
        self.items = {}

    def submit(self, item, tag=0):
        i = lib().mps_batch_submit(self.h, ctypes.byref(item), tag)
        if i < 0:
            raise RuntimeError("batch submit failed")
        self.items[i] = item
        return i

    def poll(self, cap=1024):
        arr = (ItemDone * cap)()
        n = ctypes.c_size_t()
        lib().mps_batch_poll(self.h, arr, cap, ctypes.byref(n))
        return [arr[i].as_dict() for i in range(n.value)]

    def wait_all(self):
        rc = lib().mps_batch_wait_all(self.h)
        if rc != MPS_OK:
            raise RuntimeError(f"batch engine: {self.last_error()}")

    def relations(self, item_id):
        cap = 1 << 12
        while True:
            rels = (Relation * cap)()
            n = ctypes.c_size_t()
            rc = lib().mps_batch_relations(self.h, item_id, rels, cap, ctypes.byref(n))
            if rc == MPS_CAPACITY:
                cap = n.value + 16
                continue
            if rc != MPS_OK:
                raise RuntimeError(f"relations of item {item_id}: status {rc}")
            return [(rels[i].p, rels[i].M, rels[i].q) for i in range(n.value)]

    def clock(self):
        return lib().mps_batch_clock(self.h)

    def stats(self):
        s = BatchStats()
        lib().mps_batch_get_stats(self.h, ctypes.byref(s))
        return s.as_dict()

    def last_error(self):
        return lib().mps_batch_last_error(self.h).decode() if self.h else "no engine"

    def close(self):
        if self.h:
            lib().mps_batch_destroy(self.h)
            self.h = None


def results_json(item, rels, timestamp):
    arr = (Relation * max(1, len(rels)))()
    for i, (p, M, q) in enumerate(rels):
        arr[i] = Relation(p, M, q)
    ts = timestamp.encode()
    need = lib().mps_results_json(ctypes.byref(item), arr, len(rels), ts, None, 0)
    buf = ctypes.create_string_buffer(need + 1)
    lib().mps_results_json(ctypes.byref(item), arr, len(rels), ts, buf, need + 1)
    return buf.value.decode()


def is_prime64(n):
    return bool(lib().mps_is_prime64(n))


def multiplicative_order(b, p):
    return lib().mps_multiplicative_order(b, p)


def run_bsgs(p, item):
    r = Relation()
    if lib().mps_run_bsgs(p, ctypes.byref(item), ctypes.byref(r)):
        return (r.p, r.M, r.q)
    return None
