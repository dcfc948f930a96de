// The drop-in as the reference's host drives it.  A host with the reference's WorkItem
// (Types.hpp:14-26: 128-bit b, c, P0, P1, N0, N1; 64-bit cls_mod, cls_res) includes the adapter
// include/mps_gpu_siever.hpp, unchanged, and runs process_chunk's GPU branches
// (Worker.hpp:195-300) on chunks of the item (enqueue_prime_chunks, Worker.hpp:174-193):
//   sieve  (GPU_sieve=1, Worker.hpp:216-228): sieve_filter(lo, hi) per chunk, then sort, dedup,
//          range check and is_prime64 on what it returns, then run_bsgs on each prime;
//   filter (GPU_sieve=0, Worker.hpp:245-257): the host sieves the chunk itself, filter() returns
//          the indices to keep, run_bsgs on those primes.
// The host's own CPU functions (segmented_sieve, is_prime64, run_bsgs) are the oracle's
// restatements of the reference's (liboracle.so): the only part that is not the reference's host
// is the GPU worker behind the adapter.  GPU_threads feeder threads share one GpuSiever, as
// run_worker_pool does (Worker.hpp:340).
//
// Usage: dropin_check sieve|filter b c P0 P1 N0 N1 mod res [chunk_ints] [threads]
// Prints the relations as "p, M, q" lines sorted by p, then "chunks <n> declined <d>" on stderr.
// Exit status 0, or 1 when the device is missing or a chunk was declined.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../oracle/mps_oracle.h"
#include "mps_gpu_siever.hpp"

namespace {

typedef __int128 Int128;

// Types.hpp:14-26 with __int128 for boost::multiprecision::int128_t (no Boost here).
struct WorkItem {
    Int128 b{};
    Int128 c{};
    Int128 P0{};
    Int128 P0_original{};
    Int128 P1{};
    Int128 N0{};
    Int128 N1{};
    std::uint64_t relations_restored{};
    std::uint64_t cls_mod{1};
    std::uint64_t cls_res{0};
};

struct Chunk {
    uint64_t lo, hi;
};

or_item oracle_item(const WorkItem& w) {
    or_item o;
    std::memset(&o, 0, sizeof(o));
    o.b = (int64_t)w.b;
    o.c = (int64_t)w.c;
    o.P0 = (uint64_t)w.P0;
    o.P1 = (uint64_t)w.P1;
    o.N0 = (int64_t)w.N0;
    o.N1 = (int64_t)w.N1;
    o.cls_mod = w.cls_mod;
    o.cls_res = w.cls_res;
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 10) {
        std::fprintf(stderr, "usage: %s sieve|filter b c P0 P1 N0 N1 mod res [chunk_ints] [threads]\n", argv[0]);
        return 2;
    }
    const bool gpu_sieve = std::strcmp(argv[1], "sieve") == 0;
    WorkItem item;
    item.b = std::strtoll(argv[2], nullptr, 10);
    item.c = std::strtoll(argv[3], nullptr, 10);
    item.P0 = item.P0_original = std::strtoull(argv[4], nullptr, 10);
    item.P1 = std::strtoull(argv[5], nullptr, 10);
    item.N0 = std::strtoll(argv[6], nullptr, 10);
    item.N1 = std::strtoll(argv[7], nullptr, 10);
    item.cls_mod = std::strtoull(argv[8], nullptr, 10);
    item.cls_res = std::strtoull(argv[9], nullptr, 10);
    const uint64_t chunk_ints = argc > 10 ? std::strtoull(argv[10], nullptr, 10) : (1ull << 26);
    const int threads = argc > 11 ? std::atoi(argv[11]) : 2;

    mps::GpuSiever gpu;  // Siever.cpp:116-125
    if (!gpu.init(0, "auto")) {
        std::fprintf(stderr, "%s\n", gpu.banner().c_str());
        return 1;
    }
    std::fprintf(stderr, "%s\n", gpu.banner().c_str());
    gpu.set_classes(true);

    std::vector<Chunk> chunks;  // enqueue_prime_chunks
    for (uint64_t cur = (uint64_t)item.P0; cur < (uint64_t)item.P1;) {
        const uint64_t step = std::min<uint64_t>(chunk_ints, (uint64_t)item.P1 - cur);
        chunks.push_back(Chunk{cur, cur + step});
        cur += step;
    }
    const or_item oi = oracle_item(item);
    std::atomic<size_t> next{0};
    std::atomic<int> declined{0};
    std::mutex mu;
    std::vector<or_relation> rels;
    auto feeder = [&]() {
        for (size_t k; (k = next.fetch_add(1)) < chunks.size();) {
            const uint64_t lo = chunks[k].lo, hi = chunks[k].hi;
            std::vector<uint64_t> primes;
            if (gpu_sieve) {
                std::vector<uint64_t> flagged;
                if (!gpu.sieve_filter(lo, hi, item, flagged)) {  // GPU_strict=1: a decline stops the run
                    std::fprintf(stderr, "chunk [%llu, %llu) declined: %s\n", (unsigned long long)lo,
                                 (unsigned long long)hi, mps_gpu_last_error(gpu.handle()));
                    declined.fetch_add(1);
                    continue;
                }
                std::sort(flagged.begin(), flagged.end());
                flagged.erase(std::unique(flagged.begin(), flagged.end()), flagged.end());
                for (const uint64_t p : flagged)
                    if (p >= lo && p < hi && or_is_prime64(p)) primes.push_back(p);
            } else {
                const size_t n = or_sieve_primes(lo, hi, nullptr, 0);  // segmented_sieve
                primes.resize(n);
                if (n) or_sieve_primes(lo, hi, primes.data(), n);
                std::vector<std::size_t> todo;
                if (!gpu.filter(primes, item, todo)) {
                    std::fprintf(stderr, "chunk [%llu, %llu) declined: %s\n", (unsigned long long)lo,
                                 (unsigned long long)hi, mps_gpu_last_error(gpu.handle()));
                    declined.fetch_add(1);
                    continue;
                }
                std::vector<uint64_t> kept;
                for (const std::size_t i : todo) kept.push_back(primes[i]);
                primes.swap(kept);
            }
            std::vector<or_relation> mine;
            for (const uint64_t p : primes) {
                or_relation r;
                if (or_run_bsgs(p, &oi, &r)) mine.push_back(r);
            }
            std::lock_guard<std::mutex> l(mu);
            rels.insert(rels.end(), mine.begin(), mine.end());
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(feeder);
    for (auto& t : pool) t.join();
    std::sort(rels.begin(), rels.end(), [](const or_relation& a, const or_relation& b) { return a.p < b.p; });
    for (const or_relation& r : rels)
        std::printf("%llu, %llu, %llu\n", (unsigned long long)r.p, (unsigned long long)r.M, (unsigned long long)r.q);
    std::fprintf(stderr, "chunks %zu declined %d\n", chunks.size(), declined.load());
    return declined.load() ? 1 : 0;
}
