/*
 * mps_oracle.h -- CPU oracle for the Mass Prime Siever GPU worker.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the reference's CPU path
 * (Math.hpp, Sieve.hpp, the CPU branch of Worker.hpp, Results.hpp) and the semantics of the
 * reference's device stages (Gpu.hpp psieve / prep_core / prep_q_core) so that tests, the smoke
 * check and the benchmark's CPU baseline can check the HIP worker against it.  Nothing in the
 * product links, loads or calls this library.
 *
 * Parity pins (every function is checked against something it does not share code with):
 *   or_mont_mul          closed form a*b*2^-64 mod n in Python ints (tests/test_oracle.py)
 *   or_is_prime64        sympy.isprime; strong pseudoprimes 2047, 3825123056546413051
 *   or_multiplicative_order  sympy.n_order
 *   or_run_bsgs          brute force pow(b, n, p) over every n of small items
 *   or_sieve_primes      sympy.primerange
 *   or_sieve_survivors   trial division by the odd primes <= 65521
 *   or_prep / or_prep_q  Euler criterion / g-th power residuosity with pow() in Python ints,
 *                        and the no-false-negative property against brute force
 *   or_results_json      hand-checked fixture bytes (tests/golden/) and a Python restatement
 */
#ifndef MPS_ORACLE_H
#define MPS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One worktodo item (Types.hpp:14-26).  All values on this path are below 2^63. */
typedef struct or_item {
    int64_t b;
    int64_t c;
    uint64_t P0;
    uint64_t P1;
    int64_t N0;
    int64_t N1;
    uint64_t cls_mod; /* 1 (or 0) = no class restriction */
    uint64_t cls_res;
} or_item;

typedef struct or_relation {
    uint64_t p;
    uint64_t M;
    uint64_t q;
} or_relation;

/* Result of the device-stage restatement of prep_core (Gpu.hpp:231-308). */
typedef struct or_entry {
    uint64_t jcount; /* giant range j in [0, jcount) */
    uint64_t m;      /* baby steps */
    uint32_t stride; /* 1 or 2 */
    uint32_t parity; /* 0 or 1 */
} or_entry;

/* Result of the device-stage restatement of prep_q_core (Gpu.hpp:478-589). */
typedef struct or_entry_q {
    uint32_t qp;    /* lcm(cm, g) */
    uint32_t ncls;  /* qp / cm */
    uint32_t mask;  /* classes that keep a possible solution */
    uint32_t s;     /* popcount(mask) */
    uint64_t m;     /* baby steps (<= 1024) */
    uint64_t Gc;    /* giant steps per class */
} or_entry_q;

uint64_t or_mont_mul(uint64_t a, uint64_t b, uint64_t n);
uint64_t or_powmod(uint64_t base, uint64_t e, uint64_t mod);
int or_is_prime64(uint64_t n);
uint64_t or_multiplicative_order(uint64_t base, uint64_t prime);
int or_jacobi(uint64_t a, uint64_t n);

/* run_bsgs (Math.hpp:377-486) with its run_bsgs_reference fallback (Math.hpp:488-542). */
int or_run_bsgs(uint64_t prime, const or_item* item, or_relation* out);

/* Primes in [lo, hi) by a segmented sieve (Sieve.hpp:110-185).  Returns the count written;
 * when out is NULL only counts.  Requires hi < 2^63. */
size_t or_sieve_primes(uint64_t lo, uint64_t hi, uint64_t* out, size_t cap);

/* psieve semantics (Gpu.hpp:385-455): odd x in [lo|1, hi) with no odd prime factor q <= 65521
 * other than x itself.  Requires lo > 3. */
size_t or_sieve_survivors(uint64_t lo, uint64_t hi, uint64_t* out, size_t cap);

/* gpu_classes (Math.hpp:324-337).  Returns 1 when the class split is used. */
int or_gpu_classes(const or_item* item, uint64_t* Q, uint32_t* classes, uint32_t* nclasses);

/* prep_core / prep_q_core decisions for one number x.  Return 0 = dropped, 1 = entry, 2 = CPU. */
int or_prep(uint64_t x, const or_item* item, int need_prime, or_entry* out);
int or_prep_q(uint64_t x, const or_item* item, uint64_t Qmax, int need_prime, or_entry_q* out);

/* CPU branch of process_chunk over [lo, hi) (Worker.hpp:243-280), split into chunks of
 * chunk_ints integers and run on nthreads threads.  Relations are returned sorted by p.
 * Returns the number of relations, or (size_t)-1 when cap is too small. */
size_t or_process_range(const or_item* item, uint64_t lo, uint64_t hi, uint64_t chunk_ints, int nthreads,
                        or_relation* out, size_t cap, uint64_t* primes_seen);

/* apply_relations (Worker.hpp:527-559) + one results.json object (Results.hpp:72-107), compact
 * like boost::json::serialize.  Relations are applied in the given order.  Returns the length
 * written (without the terminating NUL), or the length needed when cap is too small. */
size_t or_results_json(const or_item* item, const or_relation* rels, size_t nrel, const char* timestamp,
                       char* buf, size_t cap);

#ifdef __cplusplus
}
#endif

#endif
