// CPU stand-in for the device side of the batch engine (test infrastructure only).
//
// It provides the HIP runtime calls of tests/mock_hip/hip/hip_runtime.h and the device-facing
// helpers the engine links against (plan_item, pass_field, build_pass_graph, ...), so
// that the real scheduler in mass-prime-siever_amd/csrc/mps_batch.hip runs on the CPU.
//
// A simulated pass reads its descriptor when it runs (as the device does), checks that the
// item's class-size table reached "device" memory, emits the first few odd numbers of the pass as
// records the host must decide (M = MPS_HOST_DECIDES), sleeps for the configured pass time and
// copies counters and records back.  The host then runs the exact is_prime64 + run_bsgs on them,
// so every item's relations are a known function of the chunks that ran.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>

#include "mps_gpu.h"
#include "mps_kernels.h"
#include "mps_worker_internal.h"

namespace {

std::atomic<int> g_fail_at{0};       // 1-based pass number that fails (0: none)
std::atomic<int> g_pass_us{200};     // simulated device time per pass
std::atomic<int> g_passes{0};
std::atomic<int> g_running{0};
std::atomic<int> g_max_running{0};
std::atomic<int> g_table_mismatch{0};
constexpr uint32_t kRecordsPerPass = 8;

}  // namespace

struct MockStream {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> q;
    bool busy = false, quit = false;
    hipError_t err = hipSuccess;
    std::thread th;

    MockStream() {
        th = std::thread([this] {
            std::unique_lock<std::mutex> l(mu);
            for (;;) {
                cv.wait(l, [this] { return quit || !q.empty(); });
                if (q.empty()) return;
                auto op = std::move(q.front());
                q.pop_front();
                busy = true;
                const bool skip = err != hipSuccess;  // errors are sticky, like a faulted queue
                l.unlock();
                if (!skip) op();
                l.lock();
                busy = false;
                cv.notify_all();
            }
        });
    }
    void push(std::function<void()> f) {
        std::lock_guard<std::mutex> l(mu);
        q.push_back(std::move(f));
        cv.notify_all();
    }
};

struct MockGraph {
    mps::PassChain c;
};
struct MockGraphExec {
    mps::PassChain c;
};

namespace {

void run_pass(MockStream* s, const mps::PassChain& c) {
    std::memcpy(c.d_desc, c.h_desc, sizeof(mps::PassDesc));
    const int n = ++g_passes;
    if (n == g_fail_at.load()) {
        std::lock_guard<std::mutex> l(s->mu);
        s->err = hipErrorLaunchFailure;
        return;
    }
    const int r = ++g_running;
    int m = g_max_running.load();
    while (r > m && !g_max_running.compare_exchange_weak(m, r)) {
    }
    const mps::PassDesc* d = c.d_desc;
    std::memset(c.pb.cnt, 0, sizeof(mps::Counters));
    // the item's class table must be in device memory, tagged with the item's b and c
    if (d->ip.csz == nullptr || d->ip.csz[0].x != (uint32_t)d->ip.b || d->ip.csz[0].y != (uint32_t)d->ip.c) {
        ++g_table_mismatch;
    } else {
        const uint32_t k = (uint32_t)std::min<uint64_t>(d->nodd, kRecordsPerPass);
        for (uint32_t i = 0; i < k && i < c.pb.out_cap; ++i) {
            c.pb.out[2 * i] = d->lo + 2 * (uint64_t)i;
            c.pb.out[2 * i + 1] = mps::kHostDecides;
        }
        c.pb.cnt->n_out = k;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(g_pass_us.load()));
    --g_running;
    std::memcpy(c.h_cnt, c.pb.cnt, sizeof(mps::Counters));
    std::memcpy(c.h_out, c.pb.out, c.out_words * sizeof(uint64_t));
}

}  // namespace

const char* hipGetErrorString(hipError_t e) {
    return e == hipSuccess ? "no error" : e == hipErrorLaunchFailure ? "unspecified launch failure" : "mock error";
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipMallocBytes(void** p, size_t n) {
    *p = std::calloc(1, n ? n : 1);
    return *p ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMallocBytes(void** p, size_t n, unsigned) { return hipMallocBytes(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t s) {
    s->push([=] { std::memcpy(dst, src, n); });
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = new MockStream();
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    {
        std::lock_guard<std::mutex> l(s->mu);
        s->quit = true;
        s->cv.notify_all();
    }
    s->th.join();
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    std::unique_lock<std::mutex> l(s->mu);
    s->cv.wait(l, [s] { return s->q.empty() && !s->busy; });
    return s->err;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new MockGraphExec{g->c};
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    const mps::PassChain c = e->c;
    s->push([s, c] { run_pass(s, c); });
    return hipSuccess;
}

namespace mps {

bool plan_item(const mps_gpu*, const mps_item* item, uint64_t, bool, Plan& pl, std::string& why) {
    if (item->N1 - item->N0 > 1000000000000ll) {
        why = "BSGS table would exceed the LDS budget";
        return false;
    }
    pl.ip = ItemParams{};
    pl.ip.b = item->b;
    pl.ip.c = item->c;
    pl.ip.qmode = item->cls_mod > 1 ? 1 : 0;
    pl.m_max = 64;
    pl.tg = TableGeom{4, 16, 16, item->cls_mod > 1 ? 2u : 4u};
    pl.csz.fill(uint2{0, 0});
    pl.csz[0] = uint2{(uint32_t)item->b, (uint32_t)item->c};
    return true;
}

FieldKind pass_field(const mps_gpu*, uint64_t hi) { return hi <= (1ull << 50) ? FIELD_F64 : FIELD_MONT; }

uint32_t bsgs_persistent_grid(FieldKind, const TableGeom&, int cus) { return (uint32_t)cus; }

hipError_t enqueue_pass_chain(hipStream_t s, const PassChain& c) {
    s->push([s, c] { run_pass(s, c); });
    return hipSuccess;
}

hipError_t build_pass_graph(const PassChain& c, hipGraph_t* out) {
    *out = new MockGraph{c};
    return hipSuccess;
}

}  // namespace mps

extern "C" {

int mps_gpu_ok(const mps_gpu* g) { return g != nullptr && g->ok ? 1 : 0; }

int mps_gpu_sieve_relations(mps_gpu*, uint64_t, uint64_t, const mps_item*, mps_dev_relation*, size_t, size_t* n) {
    if (n) *n = 0;
    return MPS_ERROR;
}

mps_gpu* mock_gpu_create(int num_cu) {
    mps_gpu* g = new mps_gpu();
    g->ok = true;
    g->num_cu = num_cu;
    return g;
}
void mock_gpu_destroy(mps_gpu* g) { delete g; }
void mock_configure(int fail_at, int pass_us) {
    g_fail_at = fail_at;
    g_pass_us = pass_us;
    g_passes = 0;
    g_max_running = 0;
    g_table_mismatch = 0;
}
int mock_passes() { return g_passes.load(); }
int mock_max_running() { return g_max_running.load(); }
int mock_table_mismatch() { return g_table_mismatch.load(); }
uint32_t mock_records_per_pass() { return kRecordsPerPass; }

}  // extern "C"
