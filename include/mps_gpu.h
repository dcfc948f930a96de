/*
 * mps_gpu.h -- C-ABI of the MI355X GPU worker for Mass Prime Siever.
 *
 * This is the drop-in boundary: it replaces the reference's GpuSiever chunk interface
 * (Gpu.hpp:1040-1074) and the OpenCL/CUDA host workers behind it (Gpu.hpp:654-1036,
 * Cuda.hpp:195-455).  Plain C types only: no exceptions, no C++ or torch types cross it.
 * A handle may be used from several threads.  It has two pass lanes (a HIP stream with its own
 * buffers each): up to two calls on one handle run at once, so that a second feeder thread's
 * chunk runs on the device while the first one's records come back, where the reference
 * serialises its feeder threads with a mutex (Gpu.hpp:1059, :1065); further calls wait for a lane.
 * The settings calls (mps_gpu_set_classes, mps_gpu_set_class_policy, mps_gpu_configure) and
 * mps_gpu_autotune wait until no call runs on the handle.  Results do not depend on the overlap.
 */
#ifndef MPS_GPU_H
#define MPS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The fields of a work item that the device needs (Types.hpp:14-26), converted as
 * Gpu.hpp:862-864 and :896-901 convert them: N0 < 0 is searched from 0, cls_mod 0 means 1. */
typedef struct mps_item {
    int64_t b;        /* base b of b^n + c */
    int64_t c;        /* constant c */
    int64_t N0;       /* exponents n in [N0, N1) */
    int64_t N1;
    uint64_t cls_mod; /* 0 or 1: every n; else only n == cls_res (mod cls_mod) */
    uint64_t cls_res;
} mps_item;

/* Status codes.  DECLINED: this chunk must take the caller's CPU path (reference: sieve_filter
 * returning false for a range, table or buffer limit, Gpu.hpp:858, :875, :925).  ERROR: a device
 * error; the GPU sieve stays disabled for the rest of the process (sv_fail, Gpu.hpp:982-987).
 * CAPACITY: the output buffer is too small; *out_n holds the count needed (a repeated
 * mps_gpu_sieve_filter call can need a few more where a pass filled the device's entry list: the
 * numbers past it are flagged for the host, and which ones depends on the order of the device's
 * slot reservations; mps_gpu_filter sizes its passes so that this cannot happen).  STOPPED: the caller
 * asked the run to stop (mps_run_opts.stop); the chunks that finished are reported. */
enum { MPS_OK = 0, MPS_DECLINED = 1, MPS_ERROR = 2, MPS_CAPACITY = 3, MPS_STOPPED = 4 };

typedef struct mps_gpu mps_gpu;

/* GpuSiever::init (Gpu.hpp:1042-1056).  Opens HIP device `device_index` (gfx950 only) and writes
 * a banner "GPU enabled (HIP): <name> ..." into banner (may be NULL).  Returns NULL on failure;
 * then banner holds the reason ("GPU disabled: ..."). */
mps_gpu* mps_gpu_create(int device_index, char* banner, size_t banner_cap);

/* GpuSiever::ok (Gpu.hpp:1057). */
int mps_gpu_ok(const mps_gpu* g);

/* GpuSiever::set_classes (Gpu.hpp:1063): per-prime residue classes on the sieve path (default on). */
void mps_gpu_set_classes(mps_gpu* g, int on);

/* Which items use the per-prime residue classes when classes are on (no reference counterpart;
 * the reference's rule is gpu_classes, Math.hpp:324-337):
 *   0 = as the reference: an even class modulus cm with lcm(cm, 24) <= 48 and N1 - N0 >= 4096 lcm;
 *   1 = any cm with lcm(cm, 24) <= 48 and the same span (default);
 *   2 = order mode: every item (classes on, lcm(cm, 367567200) < 2^52) gets one exponent class per
 *       prime, n == n* (mod L), from the power residues of b and -c modulo gcd(Qo, p - 1), Qo =
 *       lcm(cm, B) for the longest of B = 367567200 (2^5 3^3 5^2 7 11 13 17), then times 19, 23,
 *       29, 31, 37, with Qo below 2^52 (Pohlig-Hellman on the device); other items as under 1.
 * Relations are identical under all three; only the device work differs. */
int mps_gpu_set_class_policy(mps_gpu* g, int policy);

/* GpuSiever::sieve_filter (Gpu.hpp:1064-1067, OpenClSiever::sieve_filter Gpu.hpp:857-979).
 * Generates the primes of [lo|1, hi) on the device and writes to out[] every number it cannot
 * rule out: each prime p of the range for which some in-class n in [N0, N1) has p | b^n + c,
 * plus the rare numbers the device leaves to the host (p < 1000, p | b, an order of the baby
 * step below the baby-step count, an entry that did not fit).  The device verifies every table
 * match exactly and tests primality with the 12-base Miller-Rabin test (a kernel of its own over
 * the pass's records), so the list holds no false positives or pseudoprimes otherwise.  Order is unspecified; the caller sorts, dedups,
 * range-checks and runs the exact check, as the reference's caller does (Worker.hpp:223-227).
 * Precondition of the GPU path: hi < 2^62 and lo > 3 (Worker.hpp:209). */
int mps_gpu_sieve_filter(mps_gpu* g, uint64_t lo, uint64_t hi, const mps_item* item, uint64_t* out,
                         size_t out_cap, size_t* out_n);

/* Exact device relations of [lo|1, hi): one record per prime p of the range for which some
 * in-class n in [N0, N1) has p | b^n + c, with M = the smallest such n (the relation of
 * run_bsgs, Math.hpp:377-486, without q = ord_p(b)).  Primality is decided on the device by
 * the deterministic 12-base Miller-Rabin test.  M == MPS_HOST_DECIDES marks a number the device
 * leaves to the caller's exact check (p < 1000, p | b, an order of b^cm below the baby-step
 * count, or no room for its BSGS entry).  This extends the chunk interface (no reference
 * counterpart); sieve_filter is the same pass returning the numbers only. */
#define MPS_HOST_DECIDES 0xFFFFFFFFFFFFFFFFull
typedef struct mps_dev_relation {
    uint64_t p;
    uint64_t M;
} mps_dev_relation;

int mps_gpu_sieve_relations(mps_gpu* g, uint64_t lo, uint64_t hi, const mps_item* item, mps_dev_relation* out,
                            size_t out_cap, size_t* out_n);

/* mps_gpu_sieve_relations without allocation.  No counterpart in the reference: this is for
 * callers that cannot allocate device or pinned memory during a call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  The range [lo|1, hi) holds at most
 * MPS_NOALLOC_MAX_INTS integers (larger ranges: MPS_ERROR, nothing runs) and runs as one device
 * pass on a scratch array of about 46 MB that the library keeps per GPU for the life of the
 * process (set up at mps_gpu_create); the pass's stage counters, survivors, candidates, BSGS
 * entries and records live there.  Results are those of mps_gpu_sieve_relations on the same
 * range.  Allocation-free calls on one handle are serialised; allocation-free calls through
 * different handles on the same GPU share its scratch array and must not overlap. */
#define MPS_NOALLOC_MAX_INTS (1ull << 20)
int mps_gpu_sieve_relations_noalloc(mps_gpu* g, uint64_t lo, uint64_t hi, const mps_item* item,
                                    mps_dev_relation* out, size_t out_cap, size_t* out_n);

/* GpuSiever::filter (Gpu.hpp:1058-1061, OpenClSiever::filter Gpu.hpp:726-798): the primes are
 * given by the caller.  Writes the ascending indices i with primes[i] possibly dividing an
 * in-class b^n + c (or left to the host) to todo[]. */
int mps_gpu_filter(mps_gpu* g, const uint64_t* primes, size_t n, const mps_item* item, uint64_t* todo,
                   size_t todo_cap, size_t* n_todo);

/* Text of the last error or decline reason on this handle (never NULL). */
const char* mps_gpu_last_error(const mps_gpu* g);

/* Releases the device buffers, stream and events of the handle. */
void mps_gpu_destroy(mps_gpu* g);

/* Counters and device times accumulated over sieve_filter calls.  Times come from HIP events
 * recorded on the worker's stream around each kernel launch. */
typedef struct mps_gpu_stats {
    uint64_t passes;
    uint64_t odds;          /* odd numbers sieved */
    uint64_t survivors;     /* numbers with no prime factor <= 65521 other than themselves */
    uint64_t candidates;    /* survivors passing the residue filters */
    uint64_t entries;       /* BSGS entries (prime passed Fermat and the power test) */
    uint64_t flagged;       /* numbers written to out[] */
    uint64_t flag2;         /* of which left to the host (p < 1000, p | b, entry overflow) */
    uint64_t fermat_fail;
    uint64_t w_fermat;      /* modular products, by stage */
    uint64_t w_setup;
    uint64_t w_bsgs;        /* baby + giant steps: sum over entries of m + s * giant steps */
    double ms_sieve, ms_filter, ms_setup, ms_bsgs;  /* summed kernel durations */
    uint64_t n_bsgs_launches;
    int field;              /* field of the last pass: 0 = Montgomery, 1 = FP64 */
    int tbits;              /* log2 of the BSGS table's bucket count in the last pass */
} mps_gpu_stats;

int mps_gpu_get_stats(const mps_gpu* g, mps_gpu_stats* s);
void mps_gpu_reset_stats(mps_gpu* g);

/* Tuning and test hooks (no counterpart in the reference).
 * field: -1 = automatic (FP64 when hi <= 2^50, else Montgomery), 0 = Montgomery, 1 = FP64 where
 * the range allows it (hi <= 2^50; above that the Montgomery field, as automatic).  Any other value,
 * or an argument outside its range, is an error and leaves every setting as it was.
 * pass_odds: odd numbers per device pass (buffers are sized from it; 0 keeps the current).
 * mcap: largest baby-step count without classes, 256..4096 (0 keeps the current; default 2048).
 * waves: wavefronts that share one BSGS entry, 1, 2 or 4 (0 = chosen per item). */
int mps_gpu_configure(mps_gpu* g, int field, uint64_t pass_odds, uint32_t mcap, uint32_t waves);

/* Shape of the BSGS stage for one kind of item (no reference counterpart).  In a tuning passed to
 * mps_gpu_set_tuning a zero field keeps the handle's default. */
typedef struct mps_tuning {
    uint32_t waves;   /* wavefronts that share one BSGS entry: 1, 2 or 4 (0 = by item) */
    uint32_t mcap;    /* plain and order mode: largest baby-step count, 256..4096 (default 2048) */
    uint32_t cmax;    /* class mode: largest baby-step count, 256..2048 (default 1024) */
    uint32_t load;    /* keys per 8-slot LDS table bucket the table is sized for: 2 or 4 (default 2) */
    uint32_t policy;  /* class policy for this kind of item (see mps_gpu_set_class_policy), 0 = the
                         handle's */
    uint32_t ext;     /* class policy 2: largest medium prime r of p - 1 (19 <= r, at most the
                         first 192 primes past Qo's) whose power residue narrows a prime's exponent
                         class, in both fields; 1 = none; 0 = the handle's default: 1024 for
                         passes in the Montgomery field (p > 2^50), none in the FP64 field */
} mps_tuning;

/* Autotuner.  Items are keyed by whether classes are on, kind of class modulus, p magnitude
 * (log2 P0 in steps of 6) and exponent span per class modulus (log2 in steps of 2).  For the key
 * of `item` whose P range starts at P0, times every candidate shape (with classes on: class
 * policy 1 or 2; waves 1/2/4 x load 2/4 x the baby-step cap of the resulting mode: 512 to 4096
 * plain and order mode, 256 to 2048 class mode; shapes whose table exceeds 64 KB of LDS are
 * skipped; then, in order mode, the winner with the medium primes of p - 1 off and on) on one pass
 * over the first 2 * sample_odds integers from P0 (device
 * time of the pass chain, best of two runs), requires every candidate's records to agree with the
 * default's (the same smallest solution wherever both decided a number; a number may be left to
 * the host under one shape only) and installs the fastest for the key.  best, best_ms, default_ms
 * may be NULL. */
int mps_gpu_autotune(mps_gpu* g, const mps_item* item, uint64_t P0, uint64_t sample_odds, mps_tuning* best,
                     double* best_ms, double* default_ms);

/* The tuning the handle applies to `item` starting at P0 (*key receives its key; may be NULL). */
int mps_gpu_get_tuning(mps_gpu* g, const mps_item* item, uint64_t P0, mps_tuning* t, uint32_t* key);

/* Installs a tuning for the key of `item` starting at P0 (for instance one saved earlier). */
int mps_gpu_set_tuning(mps_gpu* g, const mps_item* item, uint64_t P0, const mps_tuning* t);

/* Forgets every installed tuning. */
void mps_gpu_clear_tuning(mps_gpu* g);

/* Measured issue rates (operations per second over the whole chip) and FP64 FMA dependent-issue
 * latency of device `device_index`; the roofline peaks of the benchmark come from here. */
typedef struct mps_valu_rates {
    double fma64_per_s;
    double mul_lo32_per_s;
    double mul_hi32_per_s;
    double mad_u64_u32_per_s;
    double add32_per_s;
    double fma64_latency_cycles;
    double clock_mhz;
    int cus;
} mps_valu_rates;

int mps_gpu_measure_valu(int device_index, mps_valu_rates* r);

/* Copies `bytes` between two device buffers `reps` times with a plain streaming kernel (k_copy)
 * and returns the achieved read + write rate in GB/s.  Under a hardware-counter profiler its
 * FETCH_SIZE and WRITE_SIZE per dispatch calibrate the counters against a known byte count. */
int mps_gpu_copy_probe(int device_index, uint64_t bytes, int reps, double* gbps);

/* Stage outputs of one pass over [lo|1, hi) for stage-level parity tests:
 *   stage 0: survivors (values), stage 1: candidates (value, info) pairs,
 *   stage 2: entries (value, m, giant steps, s | info << 32) quadruples.
 * out receives the records as u64 words; *out_n = number of records. */
int mps_gpu_debug_stage(mps_gpu* g, int stage, uint64_t lo, uint64_t hi, const mps_item* item, uint64_t* out,
                        size_t out_cap_words, size_t* out_n);

#ifdef __cplusplus
}
#endif

#endif
