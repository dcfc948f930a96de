/*
 * mps_batch.h -- C-ABI of the batched / served MI355X worker (no reference counterpart: the
 * reference runs one chunk at a time on the device, serialised by a mutex, Gpu.hpp:1059-1067).
 *
 * A batch engine owns S slots.  A slot is a HIP stream, a set of pass buffers in HBM and pinned
 * staging, with the pass chain (descriptor copy, reset, sieve, filter, setup, BSGS, result copies)
 * built once per kernel variant as a hipGraph and replayed for every pass.  Work items are
 * submitted as they arrive; an item is admitted when its device-resident state (its class-size
 * table) fits the memory budget and fewer than max_active items are in progress; the chunks of
 * all admitted items share the slots (continuous batching), so passes of different items overlap
 * on the device while the host part of finished passes (q = ord_p(b)) runs on the slot threads.
 * Relations are exactly those of mps_run_item for the same item.
 */
#ifndef MPS_BATCH_H
#define MPS_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "mps_gpu.h"
#include "mps_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mps_batch mps_batch;

typedef struct mps_batch_opts {
    int slots;             /* passes in flight, 0 = 8, at most 16 */
    uint64_t pass_odds;    /* odd numbers per pass (one chunk = 2 * pass_odds integers), 0 = 2^25 */
    int use_graphs;        /* 1 = replay a captured hipGraph per pass, 0 = enqueue the kernels */
    int policy;            /* 0 = chunks in arrival order, 1 = round-robin over admitted items */
    int max_active;        /* admitted items at most, 0 = unlimited */
    uint64_t mem_budget;   /* bytes of device-resident per-item state (12544 per item), 0 = room for
                              1024 items */
} mps_batch_opts;

typedef struct mps_item_done {
    uint64_t id;
    uint64_t tag;
    int status;            /* MPS_OK, or the first error / decline of its chunks */
    uint64_t relations;
    double t_submit, t_admit, t_done;  /* seconds on the engine's clock */
    uint64_t primes_flagged;
} mps_item_done;

typedef struct mps_batch_stats {
    uint64_t passes, items_submitted, items_done, graph_launches, graphs_built;
    double busy_s;         /* time with at least one pass in flight */
    double device_s;       /* summed pass time (launch to completion) over slots */
    double host_s;         /* summed host time finishing records */
} mps_batch_stats;

/* Creates an engine on the device of g (g keeps the sieving primes; the engine has its own
 * streams and buffers).  Returns NULL on failure (reason in mps_gpu_last_error(g)). */
mps_batch* mps_batch_create(mps_gpu* g, const mps_batch_opts* opts);

/* Queues an item; returns its id (>= 0) or -1.  Never blocks on the device. */
int64_t mps_batch_submit(mps_batch* b, const mps_work_item* item, uint64_t tag);

/* Finished items since the last call (non-blocking); *n = count written. */
int mps_batch_poll(mps_batch* b, mps_item_done* done, size_t cap, size_t* n);

/* Blocks until every submitted item is finished.  After a device error every item still queued
 * or running finishes with MPS_ERROR (reason in mps_batch_last_error). */
int mps_batch_wait_all(mps_batch* b);

/* Relations of a finished item, sorted by p (*n = count; MPS_CAPACITY when cap is too small). */
int mps_batch_relations(mps_batch* b, uint64_t id, mps_relation* rels, size_t cap, size_t* n);

/* Seconds since the engine was created (the clock of mps_item_done). */
double mps_batch_clock(const mps_batch* b);

int mps_batch_get_stats(const mps_batch* b, mps_batch_stats* s);

/* Reason of the first device error of the engine ("" when none). */
const char* mps_batch_last_error(const mps_batch* b);

void mps_batch_destroy(mps_batch* b);

#ifdef __cplusplus
}
#endif

#endif
