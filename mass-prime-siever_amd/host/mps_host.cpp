// Host runtime: the GPU branch of process_chunk (Worker.hpp:195-303) driven by feeder threads
// (run_worker_pool, Worker.hpp:305-367), and the results.json object of an item
// (apply_relations Worker.hpp:527-559, append_results_json Results.hpp:72-140).
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mps_host.h"
#include "mps_host_internal.h"

namespace {

typedef __int128 i128;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ChunkResult {
    std::vector<mps_relation> rels;
    double dev_s = 0, chk_s = 0;
    uint64_t flagged = 0, prime_flagged = 0;
    int status = MPS_OK;
};

uint64_t powmod_u64(uint64_t b, uint64_t e, uint64_t m) {
    unsigned __int128 r = 1, x = b % m;
    for (; e; e >>= 1) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
    }
    return (uint64_t)r;
}

uint64_t mod_i64(int64_t v, uint64_t m) {
    const i128 r = (i128)v % (i128)m;
    return (uint64_t)(r < 0 ? r + (i128)m : r);
}

}  // namespace

namespace mps {

// The host part of one chunk [lo, hi) (Worker.hpp:223-228 and :270-280): the device returned exact
// relations (p, M) of the primes it decided and the numbers it leaves to the host; the host adds
// q = ord_p(b) to the former and runs the exact is_prime64 + run_bsgs on the latter.  Sorts recs.
// Returns MPS_ERROR if a device relation fails the direct check b^M == -c (mod p).
int finish_records(const mps_work_item* it, uint64_t lo, uint64_t hi, mps_dev_relation* recs, size_t n,
                   std::vector<mps_relation>& rels, uint64_t* prime_flagged) {
    int status = MPS_OK;
    std::sort(recs, recs + n,
              [](const mps_dev_relation& a, const mps_dev_relation& b) { return a.p < b.p || (a.p == b.p && a.M < b.M); });
    for (size_t i = 0; i < n; ++i) {
        if (i > 0 && recs[i].p == recs[i - 1].p) continue;
        const uint64_t p = recs[i].p;
        if (p < lo || p >= hi) continue;
        mps_relation rel;
        if (recs[i].M == MPS_HOST_DECIDES) {
            if (!mps_is_prime64(p)) continue;
            if (prime_flagged) ++*prime_flagged;
            if (mps_run_bsgs(p, it, &rel)) rels.push_back(rel);
            continue;
        }
        if (prime_flagged) ++*prime_flagged;
        // the device verified the equation exactly; check it once more before writing it
        if (powmod_u64(mod_i64(it->b, p), recs[i].M, p) != (p - mod_i64(it->c, p)) % p) {
            status = MPS_ERROR;
            continue;
        }
        rel.p = p;
        rel.M = recs[i].M;
        rel.q = mps_multiplicative_order(mod_i64(it->b, p), p);
        rels.push_back(rel);
    }
    return status;
}

int finish_records_parallel(const mps_work_item* it, uint64_t lo, uint64_t hi, mps_dev_relation* recs, size_t n,
                            std::vector<mps_relation>& rels, uint64_t* prime_flagged, int threads,
                            size_t min_records) {
    const size_t parts = std::min<size_t>(threads > 1 ? (size_t)threads : 1, n / std::max<size_t>(1, min_records / 2));
    if (parts <= 1 || n < min_records) return finish_records(it, lo, hi, recs, n, rels, prime_flagged);
    std::sort(recs, recs + n,
              [](const mps_dev_relation& a, const mps_dev_relation& b) { return a.p < b.p || (a.p == b.p && a.M < b.M); });
    // range k is [cut[k], cut[k + 1]); a cut never separates records of one number
    std::vector<size_t> cut(parts + 1, n);
    cut[0] = 0;
    for (size_t k = 1; k < parts; ++k) {
        size_t c = std::max(cut[k - 1], n * k / parts);
        while (c > 0 && c < n && recs[c].p == recs[c - 1].p) ++c;
        cut[k] = c;
    }
    std::vector<std::vector<mps_relation>> out(parts);
    std::vector<uint64_t> flagged(parts, 0);
    std::vector<int> status(parts, MPS_OK);
    auto run = [&](size_t k) {
        status[k] = finish_records(it, lo, hi, recs + cut[k], cut[k + 1] - cut[k], out[k], &flagged[k]);
    };
    std::vector<std::thread> th;
    for (size_t k = 1; k < parts; ++k) th.emplace_back(run, k);
    run(0);
    for (auto& t : th) t.join();
    int st = MPS_OK;
    for (size_t k = 0; k < parts; ++k) {
        rels.insert(rels.end(), out[k].begin(), out[k].end());
        if (prime_flagged) *prime_flagged += flagged[k];
        if (status[k] != MPS_OK) st = status[k];
    }
    return st;
}

int host_cpus() {
    static const int n = [] {
        cpu_set_t set;
        CPU_ZERO(&set);
        return sched_getaffinity(0, sizeof(set), &set) == 0 ? std::max(1, CPU_COUNT(&set)) : 4;
    }();
    return n;
}

}  // namespace mps

namespace {

// One chunk on the plain worker: device pass(es), then finish_records on `threads` threads when the
// chunk has many records (thousands, for small b and c over long exponent ranges, each needing
// q = ord_p(b) from a factorisation of p - 1).
void process_chunk(mps_gpu* g, const mps_work_item* it, uint64_t lo, uint64_t hi, std::vector<mps_dev_relation>& buf,
                   int threads, ChunkResult& r) {
    const mps_item di = {it->b, it->c, it->N0, it->N1, it->cls_mod, it->cls_res};
    if (!(hi < (1ull << 62) && lo > 3)) {  // gpu_range_ok (Worker.hpp:209)
        r.status = MPS_DECLINED;
        return;
    }
    size_t n = 0;
    const double t0 = now_s();
    int st = mps_gpu_sieve_relations(g, lo, hi, &di, buf.data(), buf.size(), &n);
    if (st == MPS_CAPACITY) {
        buf.resize(n + 1024);
        st = mps_gpu_sieve_relations(g, lo, hi, &di, buf.data(), buf.size(), &n);
    }
    const double t1 = now_s();
    r.dev_s += t1 - t0;
    if (st != MPS_OK) {
        r.status = st;
        return;
    }
    r.flagged += n;
    if (mps::finish_records_parallel(it, lo, hi, buf.data(), n, r.rels, &r.prime_flagged, threads, 512) != MPS_OK)
        r.status = MPS_ERROR;
    r.chk_s += now_s() - t1;
}

struct JsonOut {
    std::string s;
    void i128v(i128 v) {  // json_int (Results.hpp:57-63)
        const bool in64 = v <= (i128)INT64_MAX && v >= (i128)INT64_MIN;
        char d[48];
        int k = 0;
        const bool neg = v < 0;
        unsigned __int128 a = neg ? (unsigned __int128)(-v) : (unsigned __int128)v;
        do {
            d[k++] = (char)('0' + (int)(a % 10));
            a /= 10;
        } while (a);
        if (!in64) s += '"';
        if (neg) s += '-';
        while (k) s += d[--k];
        if (!in64) s += '"';
    }
};

i128 ceil_div(i128 a, i128 b) { return a >= 0 ? (a + b - 1) / b : a / b; }
i128 floor_div(i128 a, i128 b) { return a >= 0 ? a / b : (a - b + 1) / b; }

bool in_class(const mps_work_item* it, i128 n) {  // Types.hpp:28-31
    if (it->cls_mod <= 1) return true;
    return (uint64_t)(n % (i128)it->cls_mod) == it->cls_res;
}

}  // namespace

extern "C" {

int mps_run_item(mps_gpu* g, const mps_work_item* item, const mps_run_opts* opts, mps_relation* rels, size_t cap,
                 size_t* nrel, mps_run_stats* stats) {
    if (nrel) *nrel = 0;
    if (g == nullptr || item == nullptr || nrel == nullptr || !mps_gpu_ok(g)) return MPS_ERROR;
    const uint64_t chunk = (opts && opts->chunk_ints) ? opts->chunk_ints : (1ull << 26);
    const int feeders = (opts && opts->feeder_threads > 0) ? opts->feeder_threads : 2;
    std::vector<std::pair<uint64_t, uint64_t>> chunks;  // enqueue_prime_chunks (Worker.hpp:176-193)
    for (uint64_t c = item->P0; c < item->P1; c += std::min(chunk, item->P1 - c))
        chunks.emplace_back(c, c + std::min(chunk, item->P1 - c));
    // an item of fewer chunks than feeders is split evenly among them (chunks of at least 2^20
    // integers), so that every feeder's pass lane has work; the relations do not depend on where
    // the chunks end
    const uint64_t span = item->P1 > item->P0 ? item->P1 - item->P0 : 0;
    constexpr uint64_t kMinSplit = 1ull << 20;
    if (chunks.size() < (size_t)feeders && span >= 2 * kMinSplit) {
        const uint64_t parts = std::min<uint64_t>((uint64_t)feeders, span / kMinSplit);
        chunks.clear();
        for (uint64_t k = 0; k < parts; ++k)
            chunks.emplace_back(item->P0 + (uint64_t)((unsigned __int128)span * k / parts),
                                item->P0 + (uint64_t)((unsigned __int128)span * (k + 1) / parts));
    }
    // the host threads a feeder may use for one chunk's records: its share of the CPUs, at most 8
    const int finish_threads = std::min(8, std::max(1, mps::host_cpus() / feeders));
    std::atomic<size_t> next{0};
    std::mutex mu;
    ChunkResult total;
    const double t0 = now_s();
    const mps_chunk_cb on_chunk = opts ? opts->on_chunk : nullptr;
    void* const ctx = opts ? opts->ctx : nullptr;
    const volatile int* const stop = opts ? opts->stop : nullptr;
    std::atomic<bool> stopped{false};
    auto feeder = [&](int id) {
        std::vector<mps_dev_relation> buf(1 << 14);
        ChunkResult mine;
        for (size_t k; (k = next.fetch_add(1)) < chunks.size();) {
            if (stop && *stop) {  // no further chunk starts (Worker.hpp:324); chunk k stays undone
                stopped = true;
                break;
            }
            ChunkResult r;
            process_chunk(g, item, chunks[k].first, chunks[k].second, buf, finish_threads, r);
            if (on_chunk && r.status == MPS_OK) on_chunk(ctx, id, chunks[k].first, chunks[k].second, r.rels.data(), r.rels.size());
            mine.rels.insert(mine.rels.end(), r.rels.begin(), r.rels.end());
            mine.dev_s += r.dev_s;
            mine.chk_s += r.chk_s;
            mine.flagged += r.flagged;
            mine.prime_flagged += r.prime_flagged;
            if (r.status != MPS_OK) {
                std::lock_guard<std::mutex> l(mu);
                if (total.status == MPS_OK || r.status == MPS_ERROR) total.status = r.status;
                if (total.status == MPS_ERROR) next = chunks.size();
            }
        }
        std::lock_guard<std::mutex> l(mu);
        total.rels.insert(total.rels.end(), mine.rels.begin(), mine.rels.end());
        total.dev_s += mine.dev_s;
        total.chk_s += mine.chk_s;
        total.flagged += mine.flagged;
        total.prime_flagged += mine.prime_flagged;
    };
    std::vector<std::thread> th;
    for (int i = 1; i < feeders; ++i) th.emplace_back(feeder, i);
    feeder(0);
    for (auto& t : th) t.join();
    std::sort(total.rels.begin(), total.rels.end(),
              [](const mps_relation& a, const mps_relation& b) { return a.p < b.p; });
    const double wall = now_s() - t0;
    for (size_t i = 0; i < total.rels.size() && i < cap; ++i) rels[i] = total.rels[i];
    *nrel = total.rels.size();
    if (stats) {
        stats->wall_s = wall;
        stats->device_call_s = total.dev_s;
        stats->check_s = total.chk_s;
        stats->chunks = chunks.size();
        stats->declined = total.status == MPS_DECLINED ? 1 : 0;
        stats->flagged = total.flagged;
        stats->prime_flagged = total.prime_flagged;
        stats->relations = total.rels.size();
    }
    if (total.status != MPS_OK) return total.status;
    if (stopped) return MPS_STOPPED;
    return total.rels.size() > cap ? MPS_CAPACITY : MPS_OK;
}

size_t mps_results_json(const mps_work_item* it, const mps_relation* rels, size_t nrel, const char* timestamp,
                        char* buf, size_t cap) {
    const i128 N0 = it->N0, N1 = it->N1;
    const size_t count = N1 > N0 ? (size_t)(N1 - N0) : 0;
    // factors per n, in relation order (apply_relations, Worker.hpp:527-559)
    std::vector<uint32_t> cnt(count + 1, 0);
    std::vector<size_t> start;
    std::vector<uint64_t> fac;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            start.assign(count + 1, 0);
            size_t acc = 0;
            for (size_t i = 0; i < count; ++i) {
                start[i] = acc;
                acc += cnt[i];
                cnt[i] = 0;
            }
            start[count] = acc;
            fac.assign(acc + 1, 0);
        }
        for (size_t r = 0; r < nrel; ++r) {
            const i128 q = rels[r].q, M = rels[r].M, p = rels[r].p;
            if (q <= 0) continue;
            const i128 x0 = ceil_div(N0 - M, q), x1 = floor_div(N1 - 1 - M, q);
            for (i128 x = x0; x <= x1; ++x) {
                const i128 n = q * x + M;
                const i128 index = n - N0;
                if (index < 0 || index >= (i128)count) continue;
                i128 value = 1;  // b^n + c == p: the number is the prime itself
                bool fits = true;
                for (i128 e = 0; e < n && fits; ++e) {
                    value *= it->b;
                    if (value > p * 2) fits = false;
                }
                if (fits && value + it->c == p) continue;
                if (pass == 0)
                    cnt[(size_t)index]++;
                else
                    fac[start[(size_t)index] + cnt[(size_t)index]++] = rels[r].p;
            }
        }
    }
    JsonOut o;
    o.s.reserve(count * 4 + 256);
    o.s += "{\"Timestamp\":\"";
    o.s += timestamp ? timestamp : "";
    o.s += "\",\"b\":";
    o.i128v(it->b);
    o.s += ",\"c\":";
    o.i128v(it->c);
    o.s += ",\"P0\":";
    o.i128v((i128)it->P0_original);
    o.s += ",\"P1\":";
    o.i128v((i128)it->P1);
    o.s += ",\"N0\":";
    o.i128v(N0);
    o.s += ",\"N1\":";
    o.i128v(N1);
    o.s += ",\"prp\":[";
    bool first = true;
    for (size_t i = 0; i < count; ++i) {
        const i128 n = N0 + (i128)i;
        if (!in_class(it, n) || cnt[i] != 0) continue;
        if (!first) o.s += ',';
        first = false;
        o.i128v(n);
    }
    o.s += "],\"factored\":[";
    first = true;
    for (size_t i = 0; i < count; ++i) {
        const i128 n = N0 + (i128)i;
        if (!in_class(it, n) || cnt[i] == 0) continue;
        if (!first) o.s += ',';
        first = false;
        o.s += "{\"n\":";
        o.i128v(n);
        o.s += ",\"factors\":[";
        for (uint32_t k = 0; k < cnt[i]; ++k) {
            if (k) o.s += ',';
            o.i128v((i128)fac[start[i] + k]);  // json_uint: factors < 2^63
        }
        o.s += "]}";
    }
    o.s += "]}";
    if (buf != nullptr && cap > o.s.size()) std::memcpy(buf, o.s.c_str(), o.s.size() + 1);
    return o.s.size();
}

}  // extern "C"
