"""CPU checks of the batch engine's host logic (mass-prime-siever_amd/csrc/mps_batch.hip): the
real scheduler is compiled with g++ against a CPU stand-in device (tests/mock_hip), whose passes
emit the first odd numbers of each chunk as records for the host to decide.  Every item's
relations are then a known function of the chunks that ran, so the checks cover chunking, graph
replay of the pass descriptor, continuous batching over slots, admission by count and by
memory budget, declines, and the device-error path (every item finishes; none waits forever).
A ThreadSanitizer build of the same engine runs the scenarios of tests/mock_hip/drive_batch.cpp."""
import ctypes
import os
import shutil
import subprocess
import threading

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "mass-prime-siever_amd")
SOURCES = [os.path.join(PKG, "csrc", "mps_batch.hip"), os.path.join(HERE, "mock_hip", "mock_device.cpp"),
           os.path.join(PKG, "host", "mps_host.cpp"), os.path.join(PKG, "host", "mps_exact.cpp")]
INCLUDES = ["-I" + os.path.join(HERE, "mock_hip"), "-I" + os.path.join(ROOT, "include"),
            "-I" + os.path.join(PKG, "csrc")]
OK, DECLINED, ERROR = 0, 1, 2
HOST_DECIDES = (1 << 64) - 1
TABLE_BYTES = 49 * 32 * 8

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def compile_engine(out, extra=(), main=None):
    srcs = ["-x", "c++", SOURCES[0], "-x", "none"] + SOURCES[1:] + ([main] if main else [])
    cmd = ["g++", "-std=c++17", "-O1", "-g", *extra, *INCLUDES, *srcs, "-o", out, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


class BatchOpts(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("pass_odds", ctypes.c_uint64), ("use_graphs", ctypes.c_int),
                ("policy", ctypes.c_int), ("max_active", ctypes.c_int), ("mem_budget", ctypes.c_uint64)]


class ItemDone(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int),
                ("relations", ctypes.c_uint64), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("primes_flagged", ctypes.c_uint64)]


class BatchStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("passes", "items_submitted", "items_done", "graph_launches",
                                               "graphs_built")] + \
               [(n, ctypes.c_double) for n in ("busy_s", "device_s", "host_s")]


class WorkItem(ctypes.Structure):
    _fields_ = [("b", ctypes.c_int64), ("c", ctypes.c_int64), ("P0", ctypes.c_uint64), ("P1", ctypes.c_uint64),
                ("N0", ctypes.c_int64), ("N1", ctypes.c_int64), ("cls_mod", ctypes.c_uint64),
                ("cls_res", ctypes.c_uint64), ("P0_original", ctypes.c_uint64)]


class Relation(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint64), ("M", ctypes.c_uint64), ("q", ctypes.c_uint64)]


@pytest.fixture(scope="module")
def eng_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mock") / "libmps_mock.so")
    compile_engine(out, extra=("-fPIC", "-shared"))
    L = ctypes.CDLL(out)
    vp, u64, sz = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_size_t
    P = ctypes.POINTER
    L.mock_gpu_create.restype = vp
    L.mock_gpu_create.argtypes = [ctypes.c_int]
    L.mock_configure.argtypes = [ctypes.c_int, ctypes.c_int]
    L.mps_batch_create.restype = vp
    L.mps_batch_create.argtypes = [vp, P(BatchOpts)]
    L.mps_batch_submit.restype = ctypes.c_int64
    L.mps_batch_submit.argtypes = [vp, P(WorkItem), u64]
    L.mps_batch_poll.argtypes = [vp, P(ItemDone), sz, P(sz)]
    L.mps_batch_wait_all.argtypes = [vp]
    L.mps_batch_relations.argtypes = [vp, u64, vp, sz, P(sz)]
    L.mps_batch_get_stats.argtypes = [vp, P(BatchStats)]
    L.mps_batch_last_error.restype = ctypes.c_char_p
    L.mps_batch_last_error.argtypes = [vp]
    L.mps_batch_destroy.argtypes = [vp]
    L.mps_is_prime64.argtypes = [u64]
    L.mps_run_bsgs.argtypes = [u64, P(WorkItem), P(Relation)]
    L.gpu = L.mock_gpu_create(16)
    return L


class Engine:
    def __init__(self, L, fail_at=0, pass_us=200, **kw):
        L.mock_configure(fail_at, pass_us)
        self.L = L
        self.pass_odds = kw.get("pass_odds", 4096)
        opts = BatchOpts(kw.get("slots", 4), self.pass_odds, kw.get("use_graphs", 1), kw.get("policy", 0),
                         kw.get("max_active", 0), kw.get("mem_budget", 0))
        self.h = L.mps_batch_create(L.gpu, ctypes.byref(opts))
        assert self.h

    def submit(self, w):
        return self.L.mps_batch_submit(self.h, ctypes.byref(w), 0)

    def wait_all(self, timeout=60):
        res = {}
        th = threading.Thread(target=lambda: res.update(rc=self.L.mps_batch_wait_all(self.h)), daemon=True)
        th.start()
        th.join(timeout)
        assert not th.is_alive(), "wait_all did not return: an item was left unfinished"
        return res["rc"]

    def poll(self):
        arr = (ItemDone * 4096)()
        n = ctypes.c_size_t()
        self.L.mps_batch_poll(self.h, arr, 4096, ctypes.byref(n))
        return [arr[i] for i in range(n.value)]

    def relations(self, i):
        arr = (Relation * 4096)()
        n = ctypes.c_size_t()
        assert self.L.mps_batch_relations(self.h, i, arr, 4096, ctypes.byref(n)) == OK
        return [(arr[k].p, arr[k].M, arr[k].q) for k in range(n.value)]

    def stats(self):
        s = BatchStats()
        self.L.mps_batch_get_stats(self.h, ctypes.byref(s))
        return s

    def close(self):
        self.L.mps_batch_destroy(self.h)


def items(n, seed=1):
    import random
    rng = random.Random(seed)
    out = []
    for k in range(n):
        P0 = rng.randrange(10**6, 10**9)
        mod = rng.choice((1, 2, 6))
        out.append(WorkItem(rng.choice((2, 3, 10)), rng.choice((-25, 111, 7)), P0,
                            P0 + rng.randrange(1, 6 * 8192), 0, 100000, mod, rng.randrange(mod) if mod > 1 else 0,
                            P0))
    return out


def expected(L, w, pass_odds, per_pass=8):
    """Relations of the stand-in device: per chunk of 2 * pass_odds integers, the first odd
    numbers that are prime, through the exact host check."""
    rels = []
    chunk = 2 * pass_odds
    c = w.P0
    while c < w.P1:
        hi = min(c + chunk, w.P1)
        lo = c | 1
        nodd = (hi - lo + 1) // 2 if hi > lo else 0
        for i in range(min(nodd, per_pass)):
            p = lo + 2 * i
            if p < hi and L.mps_is_prime64(p):
                r = Relation()
                if L.mps_run_bsgs(p, ctypes.byref(w), ctypes.byref(r)):
                    rels.append((r.p, r.M, r.q))
        c = hi
    return sorted(rels)


@pytest.mark.parametrize("graphs,policy", [(1, 0), (0, 0), (1, 1)])
def test_relations_and_chunks(eng_lib, graphs, policy):
    L = eng_lib
    ws = items(30)
    e = Engine(L, use_graphs=graphs, policy=policy)
    try:
        ids = [e.submit(w) for w in ws]
        assert e.wait_all() == OK
        done = {d.id: d for d in e.poll()}
        assert sorted(done) == ids
        assert all(d.status == OK for d in done.values())
        assert L.mock_table_mismatch() == 0
        for i, w in zip(ids, ws):
            assert e.relations(i) == expected(L, w, e.pass_odds)
        st = e.stats()
        chunks = sum(-(-(w.P1 - w.P0) // (2 * e.pass_odds)) for w in ws)
        assert st.passes == chunks == L.mock_passes()
        assert st.graph_launches == (chunks if graphs else 0)
        assert st.items_done == st.items_submitted == len(ws)
    finally:
        e.close()


def test_slots_overlap(eng_lib):
    L = eng_lib
    e = Engine(L, pass_us=3000, slots=4)
    try:
        for w in items(8, seed=2):
            e.submit(w)
        assert e.wait_all() == OK
        assert L.mock_max_running() >= 2
    finally:
        e.close()


def overlap_max(done):
    return max(sum(1 for x in done if x.t_admit <= d.t_admit < x.t_done) for d in done)


@pytest.mark.parametrize("kw", [{"max_active": 2}, {"mem_budget": 2 * TABLE_BYTES}])
def test_admission_limits(eng_lib, kw):
    L = eng_lib
    e = Engine(L, pass_us=1000, **kw)
    try:
        ws = items(12, seed=3)
        ids = [e.submit(w) for w in ws]
        assert e.wait_all() == OK
        done = e.poll()
        assert len(done) == len(ws)
        assert overlap_max(done) <= 2
        assert all(d.t_submit <= d.t_admit <= d.t_done for d in done)
        for i, w in zip(ids, ws):
            assert e.relations(i) == expected(L, w, e.pass_odds)
    finally:
        e.close()


@pytest.mark.parametrize("graphs", [1, 0])
def test_device_error_finishes_every_item(eng_lib, graphs):
    """A failed pass marks the engine broken: the running items, the items still waiting for
    admission and items submitted afterwards all finish with MPS_ERROR; wait_all returns."""
    L = eng_lib
    e = Engine(L, fail_at=3, pass_us=500, max_active=2, use_graphs=graphs)
    try:
        ws = items(10, seed=4)
        for w in ws:
            e.submit(w)
        assert e.wait_all() == ERROR
        done = e.poll()
        assert len(done) == len(ws)
        assert any(d.status == ERROR for d in done)
        assert b"pass" in L.mps_batch_last_error(e.h)
        late = e.submit(ws[0])
        assert e.wait_all() == ERROR
        assert [d.status for d in e.poll() if d.id == late] == [ERROR]
    finally:
        e.close()


def test_declines(eng_lib):
    L = eng_lib
    e = Engine(L)
    try:
        a = e.submit(WorkItem(2, -25, 2, 20000, 0, 1000, 1, 0, 2))                      # range from 2
        b = e.submit(WorkItem(3, -2, (1 << 62) - 8192, (1 << 62) + 8192, 0, 1000, 1, 0, 0))  # past 2^62
        c = e.submit(WorkItem(2, -25, 10**9, 10**9, 0, 1000, 1, 0, 10**9))              # empty
        d = e.submit(WorkItem(2, -25, 10**9, 10**9 + 100, 0, 2 * 10**12, 1, 0, 10**9))   # table too large
        assert e.wait_all() == OK
        st = {x.id: x.status for x in e.poll()}
        assert st == {a: DECLINED, b: DECLINED, c: OK, d: DECLINED}
        assert e.relations(c) == []
    finally:
        e.close()


def test_thread_sanitizer_clean(tmp_path):
    exe = str(tmp_path / "drive_tsan")
    try:
        compile_engine(exe, extra=("-fsanitize=thread",), main=os.path.join(HERE, "mock_hip", "drive_batch.cpp"))
    except subprocess.CalledProcessError as e:
        pytest.skip(f"no ThreadSanitizer toolchain: {e.stderr[-200:]}")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout


def test_finish_records_parallel_matches_serial(tmp_path):
    """The batch engine finishes passes with many records on several threads
    (finish_records_parallel); the relations must be finish_records' own, for 1 to 6 threads."""
    exe = str(tmp_path / "host_parallel_check")
    compile_engine(exe, main=os.path.join(HERE, "host_parallel_check.cpp"))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout
