// Batched / served device worker: the C-ABI of include/mps_batch.h.
//
// The reference serialises one chunk at a time on its device (Gpu.hpp:1059-1067) and reads the
// stage counts back between kernels.  Here S slots run passes concurrently, each on its own HIP
// stream with its own pass buffers.  A pass never reads a count back before it ends: every stage
// kernel takes its element count from the device counters, the BSGS kernel is a persistent grid
// over the entries of the pass (the device queue of the pass), and the whole chain
//   descriptor copy -> reset -> sieve -> filter -> setup -> BSGS -> counter and record copies
// is built once per slot and kernel variant as a hipGraph of explicit nodes (no stream capture)
// and replayed for every pass.
//
// Items arrive through mps_batch_submit at any time.  An item is admitted when its
// device-resident state (its class-size table, kept in an arena allocated with the engine) fits
// the memory budget and fewer than max_active items are in progress; the chunks of every admitted item then share the slots (continuous
// batching).  While a slot's pass runs on the device, the slot threads of finished passes turn
// records into relations (q = ord_p(b)), so host and device work of different items overlap.
#include <hip/hip_runtime.h>

#include <sched.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mps_batch.h"
#include "mps_gpu.h"
#include "mps_host.h"
#include "mps_kernels.h"
#include "mps_worker_internal.h"
#include "../host/mps_host_internal.h"

using namespace mps;

namespace {

constexpr size_t kCszBytes = kCszEntries * sizeof(uint2);  // class-size table of one item
constexpr uint32_t kFastRecords = 8192;                 // records copied back inside the chain
constexpr uint32_t kDefaultTables = 1024;               // table arena without a memory budget

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct BItem {
    uint64_t id = 0, tag = 0;
    mps_work_item w{};
    mps_item di{};
    ItemParams ip{};
    uint64_t m_max = 0;
    TableGeom tg{};
    std::array<uint2, kCszEntries> csz;
    int table = -1;  // arena slot of the class-size table while admitted
    std::vector<std::pair<uint64_t, uint64_t>> chunks;
    size_t next = 0, running = 0;  // chunks handed out, chunks on a slot now
    std::vector<mps_relation> rels;
    int status = MPS_OK;
    uint64_t primes_flagged = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool admitted = false, done = false;
};

struct Task {
    BItem* it = nullptr;
    uint64_t lo = 0, hi = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    PassBuffers pb{};
    PassDesc* d_desc = nullptr;
    PassDesc* h_desc = nullptr;
    Counters* h_cnt = nullptr;
    uint64_t* h_out = nullptr;
    std::map<uint64_t, hipGraphExec_t> graphs;  // by kernel variant
    std::vector<mps_dev_relation> recs;
    std::thread th;
};

}  // namespace

struct mps_batch {
    mps_gpu* g = nullptr;
    mps_batch_opts o{};
    double t0 = 0;
    mutable std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<BItem*> pending;   // submitted, not admitted
    std::vector<BItem*> active;   // admitted, chunks left or in flight
    std::map<uint64_t, BItem*> items;
    std::deque<mps_item_done> done_q;
    // device-resident class-size tables of admitted items: an arena allocated once, filled from
    // pinned staging on a stream of its own
    uint2* d_tables = nullptr;
    uint2* h_stage = nullptr;
    hipStream_t up_stream = nullptr;
    std::vector<int> free_tables;
    uint64_t next_id = 0;
    uint64_t mem_used = 0;
    size_t rr = 0;
    bool stop = false;
    std::atomic<bool> broken{false};
    std::string err;
    std::vector<Slot> slots;
    uint32_t grid_sieve = 0, grid_stage = 0;
    std::atomic<int> in_flight{0};  // chunks on the slots (changed under mu)
    double busy_since = 0;
    mps_batch_stats st{};
};

namespace {

// First device error of the engine; later submissions and chunks finish with MPS_ERROR.
// fail_locked: with b->mu held.
bool fail_locked(mps_batch* b, hipError_t e, const std::string& what) {
    if (!b->broken) b->err = what + ": " + hipGetErrorString(e);
    b->broken = true;
    return false;
}

bool batch_fail(mps_batch* b, hipError_t e, const std::string& what) {
    std::lock_guard<std::mutex> l(b->mu);
    return fail_locked(b, e, what);
}

#define BCHK(expr, what)                                      \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return batch_fail(b, e_, what); \
    } while (0)

bool alloc_slot(mps_batch* b, Slot& s) {
    const uint64_t odds = b->o.pass_odds;
    const uint64_t ent_cap = std::max<uint64_t>(odds / 8, 4096);
    const uint64_t out_cap = record_cap(odds);
    BCHK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "slot stream");
    BCHK(hipMalloc(&s.pb.surv, odds * sizeof(uint32_t)), "surv alloc");
    BCHK(hipMalloc(&s.pb.cand, odds * sizeof(uint64_t)), "cand alloc");
    BCHK(hipMalloc(&s.pb.ent, ent_cap * kEntWords * sizeof(uint64_t)), "entry alloc");
    BCHK(hipMalloc(&s.pb.out, 2 * out_cap * sizeof(uint64_t)), "out alloc");
    BCHK(hipMalloc(&s.pb.cnt, sizeof(Counters)), "counter alloc");
    BCHK(hipMalloc(&s.d_desc, sizeof(PassDesc)), "descriptor alloc");
    BCHK(hipHostMalloc((void**)&s.h_desc, sizeof(PassDesc), hipHostMallocDefault), "pinned descriptor");
    BCHK(hipHostMalloc((void**)&s.h_cnt, sizeof(Counters), hipHostMallocDefault), "pinned counters");
    BCHK(hipHostMalloc((void**)&s.h_out, 2 * out_cap * sizeof(uint64_t), hipHostMallocDefault), "pinned out");
    s.pb.surv_cap = (uint32_t)odds;
    s.pb.cand_cap = (uint32_t)odds;
    s.pb.ent_cap = (uint32_t)ent_cap;
    s.pb.out_cap = (uint32_t)out_cap;
    s.pb.vals = nullptr;
    s.pb.need_prime = 1;
    s.pb.out_index = 0;
    return true;
}

void free_slot(Slot& s) {
    for (auto& kv : s.graphs) hipGraphExecDestroy(kv.second);
    s.graphs.clear();
    if (s.stream) hipStreamDestroy(s.stream);
    hipFree(s.pb.surv);
    hipFree(s.pb.cand);
    hipFree(s.pb.ent);
    hipFree(s.pb.out);
    hipFree(s.pb.cnt);
    hipFree(s.d_desc);
    hipHostFree(s.h_desc);
    hipHostFree(s.h_cnt);
    hipHostFree(s.h_out);
}

PassChain slot_chain(const mps_batch* b, const Slot& s, FieldKind fk, uint32_t mode, const TableGeom& tg) {
    PassChain c{};
    c.d_desc = s.d_desc;
    c.h_desc = s.h_desc;
    c.h_cnt = s.h_cnt;
    c.h_out = s.h_out;
    c.out_words = 2 * (size_t)std::min(kFastRecords, s.pb.out_cap);
    c.tab = b->g->tab;
    c.pb = s.pb;
    c.grid_sieve = b->grid_sieve;
    c.grid_stage = b->grid_stage;
    c.grid_bsgs = bsgs_persistent_grid(fk, tg, b->g->num_cu);
    c.fk = fk;
    c.mode = mode;
    c.tg = tg;
    return c;
}

// Puts one pass on the slot's stream: a replay of the variant's graph (built from explicit
// nodes on first use), or the chain's kernels one by one.
bool launch_pass(mps_batch* b, Slot& s, FieldKind fk, uint32_t mode, const TableGeom& tg) {
    if (!b->o.use_graphs) {
        BCHK(enqueue_pass_chain(s.stream, slot_chain(b, s, fk, mode, tg)), "pass launch");
        return true;
    }
    const uint64_t key = (uint64_t)fk | (uint64_t)mode << 1 | (uint64_t)tg.waves << 4 | (uint64_t)tg.nbbits << 8;
    auto it = s.graphs.find(key);
    if (it == s.graphs.end()) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        BCHK(build_pass_graph(slot_chain(b, s, fk, mode, tg), &graph), "graph build");
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        BCHK(ie, "graph instantiate");
        it = s.graphs.emplace(key, exec).first;
        std::lock_guard<std::mutex> l(b->mu);
        b->st.graphs_built++;
    }
    BCHK(hipGraphLaunch(it->second, s.stream), "graph launch");
    std::lock_guard<std::mutex> l(b->mu);
    b->st.graph_launches++;
    return true;
}

// Items leave pending in arrival order while the active count, the memory budget and the table
// arena allow.  After a device error they finish at once with MPS_ERROR.  Called with b->mu held.
void admit(mps_batch* b) {
    while (!b->pending.empty()) {
        BItem* it = b->pending.front();
        if (!b->broken) {
            if (b->o.max_active > 0 && (int)b->active.size() >= b->o.max_active) return;
            if (b->o.mem_budget && b->mem_used + kCszBytes > b->o.mem_budget && !b->active.empty()) return;
            if (b->free_tables.empty()) return;
        }
        b->pending.pop_front();
        it->admitted = true;
        it->t_admit = now_s() - b->t0;
        std::string why;
        if (b->broken) {
            it->status = MPS_ERROR;
            it->chunks.clear();
        } else {
            Plan pl;
            if (!plan_item(b->g, &it->di, it->w.P0, b->g->classes_on, pl, why)) {
                it->status = MPS_DECLINED;
                it->chunks.clear();
            } else {
                it->ip = pl.ip;
                it->m_max = pl.m_max;
                it->tg = pl.tg;
                it->csz = pl.csz;
            }
        }
        if (!it->chunks.empty()) {
            const int t = b->free_tables.back();
            uint2* dst = b->d_tables + (size_t)t * (kCszBytes / sizeof(uint2));
            std::memcpy(b->h_stage, it->csz.data(), it->csz.size() * sizeof(uint2));
            hipError_t e = hipMemcpyAsync(dst, b->h_stage, it->csz.size() * sizeof(uint2), hipMemcpyHostToDevice,
                                          b->up_stream);
            if (e == hipSuccess) e = hipStreamSynchronize(b->up_stream);
            if (e != hipSuccess) {
                fail_locked(b, e, "class table upload");
                it->status = MPS_ERROR;
                it->chunks.clear();
            } else {
                b->free_tables.pop_back();
                it->table = t;
                it->ip.csz = dst;
                b->mem_used += kCszBytes;
            }
        }
        b->active.push_back(it);
    }
}

// Called with b->mu held, once the item's last chunk is finished.
void complete(mps_batch* b, BItem* it) {
    std::sort(it->rels.begin(), it->rels.end(), [](const mps_relation& x, const mps_relation& y) { return x.p < y.p; });
    it->done = true;
    it->t_done = now_s() - b->t0;
    if (it->table >= 0) {
        b->free_tables.push_back(it->table);
        it->table = -1;
        b->mem_used -= kCszBytes;
    }
    b->active.erase(std::find(b->active.begin(), b->active.end(), it));
    mps_item_done d{};
    d.id = it->id;
    d.tag = it->tag;
    d.status = it->status;
    d.relations = it->rels.size();
    d.t_submit = it->t_submit;
    d.t_admit = it->t_admit;
    d.t_done = it->t_done;
    d.primes_flagged = it->primes_flagged;
    b->done_q.push_back(d);
    b->st.items_done++;
    b->cv_done.notify_all();
    b->cv_work.notify_all();
}

// Next chunk under the policy; admits pending items first.  Called with b->mu held.
bool next_task(mps_batch* b, Task& t) {
    admit(b);
    // items with nothing left to run (declined plans, empty ranges, errors) finish here
    for (size_t i = 0; i < b->active.size();) {
        BItem* it = b->active[i];
        if (it->next == it->chunks.size() && it->running == 0) {
            complete(b, it);
            admit(b);
            i = 0;
            continue;
        }
        ++i;
    }
    const size_t n = b->active.size();
    for (size_t k = 0; k < n; ++k) {
        const size_t i = b->o.policy == 1 ? (b->rr + k) % n : k;
        BItem* it = b->active[i];
        if (it->next < it->chunks.size()) {
            t.it = it;
            t.lo = it->chunks[it->next].first;
            t.hi = it->chunks[it->next].second;
            it->next++;
            it->running++;
            b->rr = i + 1;
            return true;
        }
    }
    return false;
}

std::string chunk_name(const Task& t) {
    char buf[96];
    std::snprintf(buf, sizeof(buf), " (item %llu, chunk at %llu)", (unsigned long long)t.it->id,
                  (unsigned long long)t.lo);
    return buf;
}

// One chunk [lo, hi) of an item on slot s: returns the chunk's status.
int run_chunk(mps_batch* b, Slot& s, const Task& t, std::vector<mps_relation>& rels, uint64_t& prime_flagged,
              double& dev_s, double& host_s) {
    BItem* it = t.it;
    if (!(t.hi < (1ull << 62) && t.lo > 3)) return MPS_DECLINED;  // gpu_range_ok (Worker.hpp:209)
    const uint64_t lo = t.lo | 1;
    if (lo >= t.hi) return MPS_OK;
    const uint64_t count = (t.hi - lo + 1) / 2;
    if (count > b->o.pass_odds) return MPS_ERROR;  // chunks are cut to one pass
    const FieldKind fk = pass_field(b->g, t.hi);
    s.h_desc->lo = lo;
    s.h_desc->nodd = count;
    s.h_desc->ip = it->ip;
    s.h_desc->tg = it->tg;
    const double t0 = now_s();
    if (!launch_pass(b, s, fk, setup_variant(it->ip, fk), it->tg)) return MPS_ERROR;
    hipError_t e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) return batch_fail(b, e, "pass" + chunk_name(t)), MPS_ERROR;
    const uint32_t n_out = s.h_cnt->n_out;
    if (n_out > s.pb.out_cap) return MPS_DECLINED;  // the plain worker declines the same chunk
    if (n_out > kFastRecords) {
        e = hipMemcpyAsync(s.h_out + 2 * kFastRecords, s.pb.out + 2 * kFastRecords,
                           2 * sizeof(uint64_t) * (n_out - kFastRecords), hipMemcpyDeviceToHost, s.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
        if (e != hipSuccess) return batch_fail(b, e, "record read" + chunk_name(t)), MPS_ERROR;
    }
    const double t1 = now_s();
    dev_s += t1 - t0;
    s.recs.resize(n_out);
    for (uint32_t i = 0; i < n_out; ++i) s.recs[i] = mps_dev_relation{s.h_out[2 * i], s.h_out[2 * i + 1]};
    // a pass with many relations (small b and c over long exponent ranges: thousands, each with
    // q = ord_p(b) from a factorisation of p - 1) is finished on several threads: the CPUs shared
    // among the slots that are running chunks, at least 4 and at most 12
    const int threads = std::min(12, std::max(4, host_cpus() / std::max(1, b->in_flight.load())));
    const int st = finish_records_parallel(&it->w, lo, t.hi, s.recs.data(), n_out, rels, &prime_flagged, threads, 512);
    host_s += now_s() - t1;
    return st;
}

void slot_loop(mps_batch* b, int si) {
    Slot& s = b->slots[si];
    hipSetDevice(b->g->device);
    for (;;) {
        Task t;
        {
            std::unique_lock<std::mutex> l(b->mu);
            b->cv_work.wait(l, [&] { return b->stop || next_task(b, t); });
            if (t.it == nullptr) return;  // stopping with nothing to run
            if (b->in_flight++ == 0) b->busy_since = now_s();
        }
        std::vector<mps_relation> rels;
        uint64_t pf = 0;
        double dev_s = 0, host_s = 0;
        const int st = b->broken ? MPS_ERROR : run_chunk(b, s, t, rels, pf, dev_s, host_s);
        std::lock_guard<std::mutex> l(b->mu);
        if (--b->in_flight == 0) b->st.busy_s += now_s() - b->busy_since;
        b->st.passes++;
        b->st.device_s += dev_s;
        b->st.host_s += host_s;
        BItem* it = t.it;
        it->rels.insert(it->rels.end(), rels.begin(), rels.end());
        it->primes_flagged += pf;
        if (st != MPS_OK && (it->status == MPS_OK || st == MPS_ERROR)) it->status = st;
        if (st == MPS_ERROR) it->next = it->chunks.size();  // the item's other chunks are not run
        it->running--;
        if (it->next == it->chunks.size() && it->running == 0) complete(b, it);
        b->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

mps_batch* mps_batch_create(mps_gpu* g, const mps_batch_opts* opts) {
    if (g == nullptr || !mps_gpu_ok(g)) return nullptr;
    mps_batch* b = new mps_batch();
    b->g = g;
    if (opts) b->o = *opts;
    if (b->o.slots <= 0) b->o.slots = 8;
    if (b->o.slots > 16) b->o.slots = 16;
    if (b->o.pass_odds == 0) b->o.pass_odds = 1ull << 25;
    if (b->o.pass_odds < 4096 || b->o.pass_odds > (1ull << 30)) {
        {
            std::lock_guard<std::mutex> l(g->st_mu);
            g->err = "batch pass_odds outside [4096, 2^30]";
        }
        delete b;
        return nullptr;
    }
    b->grid_sieve = (uint32_t)((b->o.pass_odds + kSegOdds - 1) / kSegOdds);
    b->grid_stage = 8u * (uint32_t)std::max(1, g->num_cu);
    b->t0 = now_s();
    if (hipSetDevice(g->device) != hipSuccess) {
        delete b;
        return nullptr;
    }
    // table arena: the memory budget's worth of class-size tables (at least one), else a default
    uint64_t ntab = b->o.mem_budget ? std::max<uint64_t>(1, b->o.mem_budget / kCszBytes) : kDefaultTables;
    ntab = std::min<uint64_t>(ntab, 1u << 16);
    bool ok = hipStreamCreateWithFlags(&b->up_stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&b->d_tables, ntab * kCszBytes) == hipSuccess &&
              hipHostMalloc((void**)&b->h_stage, kCszBytes, hipHostMallocDefault) == hipSuccess;
    for (uint64_t t = ntab; t-- > 0;) b->free_tables.push_back((int)t);
    b->slots.resize(b->o.slots);
    for (size_t i = 0; ok && i < b->slots.size(); ++i) ok = alloc_slot(b, b->slots[i]);
    if (!ok) {
        {
            std::lock_guard<std::mutex> l(g->st_mu);
            g->err = b->err.empty() ? std::string("batch engine: device allocation failed") : b->err;
        }
        for (auto& x : b->slots) free_slot(x);
        if (b->up_stream) hipStreamDestroy(b->up_stream);
        hipFree(b->d_tables);
        hipHostFree(b->h_stage);
        delete b;
        return nullptr;
    }
    for (int i = 0; i < b->o.slots; ++i) b->slots[i].th = std::thread(slot_loop, b, i);
    return b;
}

int64_t mps_batch_submit(mps_batch* b, const mps_work_item* item, uint64_t tag) {
    if (b == nullptr || item == nullptr) return -1;
    BItem* it = new BItem();
    it->w = *item;
    it->di = mps_item{item->b, item->c, item->N0, item->N1, item->cls_mod, item->cls_res};
    it->tag = tag;
    // chunks of one pass each (enqueue_prime_chunks, Worker.hpp:176-193, with the pass as chunk)
    const uint64_t chunk = 2 * b->o.pass_odds;
    for (uint64_t c = item->P0; c < item->P1; c += std::min(chunk, item->P1 - c))
        it->chunks.emplace_back(c, c + std::min(chunk, item->P1 - c));
    std::lock_guard<std::mutex> l(b->mu);
    it->id = b->next_id++;
    it->t_submit = now_s() - b->t0;
    b->items[it->id] = it;
    b->pending.push_back(it);
    b->st.items_submitted++;
    b->cv_work.notify_all();
    return (int64_t)it->id;
}

int mps_batch_poll(mps_batch* b, mps_item_done* done, size_t cap, size_t* n) {
    if (n) *n = 0;
    if (b == nullptr || n == nullptr) return MPS_ERROR;
    std::lock_guard<std::mutex> l(b->mu);
    size_t k = 0;
    while (k < cap && !b->done_q.empty()) {
        done[k++] = b->done_q.front();
        b->done_q.pop_front();
    }
    *n = k;
    return b->broken ? MPS_ERROR : MPS_OK;
}

int mps_batch_wait_all(mps_batch* b) {
    if (b == nullptr) return MPS_ERROR;
    std::unique_lock<std::mutex> l(b->mu);
    b->cv_work.notify_all();
    b->cv_done.wait(l, [&] { return b->st.items_done == b->st.items_submitted; });
    return b->broken ? MPS_ERROR : MPS_OK;
}

int mps_batch_relations(mps_batch* b, uint64_t id, mps_relation* rels, size_t cap, size_t* n) {
    if (n) *n = 0;
    if (b == nullptr || n == nullptr) return MPS_ERROR;
    std::lock_guard<std::mutex> l(b->mu);
    auto f = b->items.find(id);
    if (f == b->items.end() || !f->second->done) return MPS_ERROR;
    const std::vector<mps_relation>& r = f->second->rels;
    for (size_t i = 0; i < r.size() && i < cap && rels; ++i) rels[i] = r[i];
    *n = r.size();
    return r.size() > cap ? MPS_CAPACITY : MPS_OK;
}

double mps_batch_clock(const mps_batch* b) { return b ? now_s() - b->t0 : 0.0; }

int mps_batch_get_stats(const mps_batch* b, mps_batch_stats* s) {
    if (b == nullptr || s == nullptr) return MPS_ERROR;
    std::lock_guard<std::mutex> l(b->mu);
    *s = b->st;
    return MPS_OK;
}

const char* mps_batch_last_error(const mps_batch* b) { return b ? b->err.c_str() : "no engine"; }

void mps_batch_destroy(mps_batch* b) {
    if (b == nullptr) return;
    {
        std::lock_guard<std::mutex> l(b->mu);
        b->stop = true;
        b->cv_work.notify_all();
    }
    for (auto& s : b->slots)
        if (s.th.joinable()) s.th.join();
    hipSetDevice(b->g->device);
    for (auto& s : b->slots) free_slot(s);
    for (auto& kv : b->items) delete kv.second;
    if (b->up_stream) hipStreamDestroy(b->up_stream);
    hipFree(b->d_tables);
    hipHostFree(b->h_stage);
    delete b;
}

}  // extern "C"
