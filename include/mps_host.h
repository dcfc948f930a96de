/*
 * mps_host.h -- C-ABI of the host runtime around the MI355X GPU worker.
 *
 * It is the caller side of the chunk interface: the GPU branch of process_chunk
 * (Worker.hpp:195-303) with the exact re-check of every flagged number (is_prime64, run_bsgs,
 * multiplicative_order: Math.hpp:209-310, :377-542), the item loop of Siever.cpp:138-184 and the
 * relations.p / results.json formats (Worker.hpp:287-300, :469-559; Results.hpp:72-140).
 * There is no CPU sieve here: a chunk the GPU declines is reported, not sieved on the host
 * (the reference's GPU_strict=1 behaviour, Worker.hpp:229-235).
 */
#ifndef MPS_HOST_H
#define MPS_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "mps_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* WorkItem (Types.hpp:14-26); every value on this path is below 2^63. */
typedef struct mps_work_item {
    int64_t b;
    int64_t c;
    uint64_t P0;
    uint64_t P1;
    int64_t N0;
    int64_t N1;
    uint64_t cls_mod; /* 1 when the worktodo line has no class */
    uint64_t cls_res;
    uint64_t P0_original; /* written as "P0" in results.json */
} mps_work_item;

/* Relation (Math.hpp:360-364): p divides b^n + c for every n == M (mod q), M the smallest
 * in-class solution n >= N0, q = ord_p(b). */
typedef struct mps_relation {
    uint64_t p;
    uint64_t M;
    uint64_t q;
} mps_relation;

/* Called once per finished chunk [lo, hi) with its relations (sorted by p), from the feeder
 * thread that ran it (feeder = 0 .. feeder_threads - 1): the hook of the reference's per-worker
 * core_<id>.p files (Worker.hpp:294-297), which checkpointing relies on. */
typedef void (*mps_chunk_cb)(void* ctx, int feeder, uint64_t lo, uint64_t hi, const mps_relation* rels, size_t n);

typedef struct mps_run_opts {
    uint64_t chunk_ints; /* integers per chunk (Worker.hpp:152-157); 0 = 2^26 */
    int feeder_threads;  /* GPU_threads (Config.hpp:23): chunks in flight; 0 = 2 */
    mps_chunk_cb on_chunk;       /* may be NULL */
    void* ctx;
    const volatile int* stop;    /* when non-NULL and *stop != 0, no further chunk starts
                                    (shutdown_requested, Worker.hpp:324) and the run returns
                                    MPS_STOPPED */
} mps_run_opts;

typedef struct mps_run_stats {
    double wall_s;        /* first chunk submitted .. last relation found */
    double device_call_s; /* summed time inside mps_gpu_sieve_filter */
    double check_s;       /* summed time of the exact host re-check */
    uint64_t chunks;
    uint64_t declined;    /* chunks the device declined (the run then fails) */
    uint64_t flagged;     /* numbers returned by the device, after dedup */
    uint64_t prime_flagged;
    uint64_t relations;
} mps_run_stats;

/* Runs [P0, P1) of one item through the GPU worker and the exact re-check.  Relations are
 * returned sorted by p.  Returns MPS_OK, MPS_DECLINED (some chunk was declined; the relations
 * of the other chunks are still returned), MPS_STOPPED (the stop flag was raised; up to cap
 * relations of the finished chunks are returned, *nrel = their count), MPS_ERROR or MPS_CAPACITY
 * (*nrel = count needed). */
int mps_run_item(mps_gpu* g, const mps_work_item* item, const mps_run_opts* opts, mps_relation* rels,
                 size_t cap, size_t* nrel, mps_run_stats* stats);

/* apply_relations (Worker.hpp:527-559) and one results.json object (Results.hpp:72-107), compact
 * like boost::json::serialize.  Returns the length of the text (without NUL); writes it when
 * cap > length. */
size_t mps_results_json(const mps_work_item* item, const mps_relation* rels, size_t nrel, const char* timestamp,
                        char* buf, size_t cap);

/* The exact host checks, exported for tests. */
int mps_is_prime64(uint64_t n);
uint64_t mps_multiplicative_order(uint64_t base, uint64_t prime);
int mps_run_bsgs(uint64_t prime, const mps_work_item* item, mps_relation* out);

#ifdef __cplusplus
}
#endif

#endif
