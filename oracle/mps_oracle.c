/*
 * mps_oracle.c -- CPU oracle for the Mass Prime Siever GPU worker.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the benchmark's CPU
 * baseline leg, always as the checker, never as the thing measured for the GPU or shipped.
 *
 * It restates the reference's CPU path in plain C (file:line citations are to the reference's
 * "Mass Prime Siever/" directory).  The reference itself cannot be compiled here: its Types.hpp
 * needs Boost.Multiprecision and Boost.LexicalCast, which the image lacks, so there is no
 * oracle/_ref build; every function below is pinned instead by checks that share no code with it
 * (listed in mps_oracle.h).
 */
#define _GNU_SOURCE
#include "mps_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

/* ---------------------------------------------------------------- modular arithmetic */

/* mulmod (Math.hpp:22-26): operands reduced first, 128-bit product, 128/64 remainder. */
static inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)(((u128)(a % m) * (u128)(b % m)) % m);
}

/* powmod_uint64 (Math.hpp:126-140). */
uint64_t or_powmod(uint64_t base, uint64_t e, uint64_t mod) {
    if (mod == 1) return 0;
    uint64_t r = 1;
    base %= mod;
    while (e) {
        if (e & 1) r = mulmod(r, base, mod);
        base = mulmod(base, base, mod);
        e >>= 1;
    }
    return r;
}

/* Montgomery (Math.hpp:29-79): odd n < 2^63, R = 2^64. */
typedef struct {
    uint64_t n, ninv, r2, one;
} mont_t;

static void mont_init(mont_t* mt, uint64_t n) {
    uint64_t inv = n; /* inverse of n mod 2^3, doubled in precision by each Newton step */
    for (int i = 0; i < 5; ++i) inv *= 2 - n * inv;
    mt->n = n;
    mt->ninv = 0 - inv;
    mt->one = (0 - n) % n;
    mt->r2 = (uint64_t)(((u128)mt->one * mt->one) % n);
}

static inline uint64_t mont_redc(const mont_t* mt, uint64_t hi, uint64_t lo) {
    const uint64_t m = lo * mt->ninv;
    const u128 mn = (u128)m * mt->n;
    const uint64_t carry = (uint64_t)(((u128)lo + (uint64_t)mn) >> 64);
    const uint64_t t = hi + (uint64_t)(mn >> 64) + carry; /* n < 2^63: no overflow */
    return t >= mt->n ? t - mt->n : t;
}

static inline uint64_t mont_mul(const mont_t* mt, uint64_t a, uint64_t b) {
    const u128 p = (u128)a * b;
    return mont_redc(mt, (uint64_t)(p >> 64), (uint64_t)p);
}

static inline uint64_t mont_to(const mont_t* mt, uint64_t a) { return mont_mul(mt, a % mt->n, mt->r2); }

static uint64_t mont_pow(const mont_t* mt, uint64_t base, uint64_t e) {
    uint64_t r = mt->one;
    while (e) {
        if (e & 1) r = mont_mul(mt, r, base);
        base = mont_mul(mt, base, base);
        e >>= 1;
    }
    return r;
}

/* Euler criterion on a Montgomery value (Math.hpp:74-78). */
static int mont_legendre(const mont_t* mt, uint64_t a_m) {
    if (a_m == 0) return 0;
    return mont_pow(mt, a_m, (mt->n - 1) >> 1) == mt->one ? 1 : -1;
}

uint64_t or_mont_mul(uint64_t a, uint64_t b, uint64_t n) {
    mont_t mt;
    mont_init(&mt, n);
    return mont_mul(&mt, a, b);
}

static uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

/* Jacobi symbol (a/n), n odd: the textbook reciprocity algorithm. */
int or_jacobi(uint64_t a, uint64_t n) {
    int s = 1;
    a %= n;
    while (a != 0) {
        while ((a & 1) == 0) {
            a >>= 1;
            if ((n & 7) == 3 || (n & 7) == 5) s = -s;
        }
        const uint64_t x = a;
        a = n;
        n = x;
        if ((a & 3) == 3 && (n & 3) == 3) s = -s;
        a %= n;
    }
    return n == 1 ? s : 0;
}

/* ---------------------------------------------------------------- primality, order */

/* is_prime64 (Math.hpp:209-251): trial division by the 12 primes 2..37, then strong probable
 * prime tests to the same 12 bases, which is exact for every n < 2^64. */
int or_is_prime64(uint64_t n) {
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i) {
        if (n == bases[i]) return 1;
        if (n % bases[i] == 0) return 0;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++s;
    }
    for (int i = 0; i < 12; ++i) {
        uint64_t x = or_powmod(bases[i] % n, d, n);
        if (x == 1 || x == n - 1) continue;
        int ok = 0;
        for (int r = 1; r < s; ++r) {
            x = mulmod(x, x, n);
            if (x == n - 1) {
                ok = 1;
                break;
            }
        }
        if (!ok) return 0;
    }
    return 1;
}

static uint64_t xs_next(uint64_t* st) { /* xorshift64* */
    uint64_t x = *st;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *st = x;
    return x * 2685821657736338717ULL;
}

/* Pollard rho with Brent's cycle detection; n odd composite.  Returns a proper divisor. */
static uint64_t rho_divisor(uint64_t n, uint64_t* rng) {
    for (;;) {
        const uint64_t c = 1 + xs_next(rng) % (n - 1);
        uint64_t y = xs_next(rng) % n, x = y, ys = y, q = 1, g = 1;
        uint64_t r = 1;
        do {
            x = y;
            for (uint64_t i = 0; i < r; ++i) y = (mulmod(y, y, n) + c) % n;
            uint64_t k = 0;
            do {
                ys = y;
                const uint64_t lim = (r - k) < 64 ? (r - k) : 64;
                for (uint64_t i = 0; i < lim; ++i) {
                    y = (mulmod(y, y, n) + c) % n;
                    q = mulmod(q, x > y ? x - y : y - x, n);
                }
                g = gcd_u64(q, n);
                k += lim;
            } while (k < r && g == 1);
            r <<= 1;
        } while (g == 1);
        if (g == n) {
            do {
                ys = (mulmod(ys, ys, n) + c) % n;
                g = gcd_u64(x > ys ? x - ys : ys - x, n);
            } while (g == 1);
        }
        if (g != n) return g;
    }
}

/* Distinct prime factors of n (ascending), n >= 1.  Returns their count. */
static int distinct_prime_factors(uint64_t n, uint64_t* out) {
    int cnt = 0;
    uint64_t stack[128];
    int sp = 0;
    for (uint64_t f = 2; f < 1000 && f * f <= n; f += (f == 2 ? 1 : 2)) {
        if (n % f == 0) {
            out[cnt++] = f;
            while (n % f == 0) n /= f;
        }
    }
    uint64_t rng = 0x2545F4914F6CDD1DULL ^ n;
    if (n > 1) stack[sp++] = n;
    while (sp > 0) {
        const uint64_t cur = stack[--sp];
        if (cur == 1) continue;
        if (or_is_prime64(cur)) {
            int dup = 0;
            for (int i = 0; i < cnt; ++i) dup |= out[i] == cur;
            if (!dup) out[cnt++] = cur;
            continue;
        }
        const uint64_t d = rho_divisor(cur, &rng);
        stack[sp++] = d;
        stack[sp++] = cur / d;
    }
    for (int i = 1; i < cnt; ++i) { /* insertion sort */
        const uint64_t v = out[i];
        int j = i - 1;
        while (j >= 0 && out[j] > v) {
            out[j + 1] = out[j];
            --j;
        }
        out[j + 1] = v;
    }
    return cnt;
}

/* multiplicative_order (Math.hpp:294-310): ord_prime(base) from the factorisation of prime-1. */
uint64_t or_multiplicative_order(uint64_t base, uint64_t prime) {
    uint64_t order = prime - 1;
    uint64_t f[64];
    const int nf = distinct_prime_factors(prime - 1, f);
    for (int i = 0; i < nf; ++i) {
        while (order % f[i] == 0 && or_powmod(base % prime, order / f[i], prime) == 1) order /= f[i];
    }
    return order;
}

/* ---------------------------------------------------------------- BSGS (exact) */

static inline uint64_t mod_i64(int64_t v, uint64_t m) { /* mod128 (Math.hpp:196-202) */
    const i128 r = (i128)v % (i128)m;
    return (uint64_t)(r < 0 ? r + (i128)m : r);
}

static inline uint64_t neg_mod_i64(int64_t v, uint64_t m) { /* mod_neg (Math.hpp:204-207) */
    const uint64_t pos = mod_i64(v, m);
    return pos == 0 ? 0 : m - pos;
}

typedef struct {
    uint64_t* keys;
    uint32_t* vals;
    uint64_t cap;
    unsigned bits;
} table_t;

static __thread table_t tls_table;

static void table_reset(table_t* t, uint64_t entries) {
    unsigned bits = 4;
    while ((1ULL << bits) < entries * 2) ++bits;
    if (t->cap < (1ULL << bits)) {
        free(t->keys);
        free(t->vals);
        t->cap = 1ULL << bits;
        t->keys = (uint64_t*)malloc(t->cap * sizeof(uint64_t));
        t->vals = (uint32_t*)malloc(t->cap * sizeof(uint32_t));
    }
    t->bits = bits;
    memset(t->keys, 0xFF, (1ULL << bits) * sizeof(uint64_t));
}

static inline uint64_t table_slot(const table_t* t, uint64_t key) {
    return (key * 0xD6E8FEB86659FD93ULL) >> (64 - t->bits);
}

/* Exact keys; keeps the first (smallest) index for a key, like BabyTable (Math.hpp:97-105). */
static void table_insert(table_t* t, uint64_t key, uint32_t val) {
    const uint64_t mask = (1ULL << t->bits) - 1;
    uint64_t h = table_slot(t, key);
    while (t->keys[h] != ~0ULL) {
        if (t->keys[h] == key) return;
        h = (h + 1) & mask;
    }
    t->keys[h] = key;
    t->vals[h] = val;
}

static int table_find(const table_t* t, uint64_t key, uint32_t* val) {
    const uint64_t mask = (1ULL << t->bits) - 1;
    uint64_t h = table_slot(t, key);
    while (t->keys[h] != ~0ULL) {
        if (t->keys[h] == key) {
            *val = t->vals[h];
            return 1;
        }
        h = (h + 1) & mask;
    }
    return 0;
}

static uint64_t isqrt_ceil_u64(uint64_t n) {
    if (n == 0) return 0;
    uint64_t r = (uint64_t)sqrtl((long double)n);
    while (r > 0 && (u128)r * r > n) --r;
    while ((u128)(r + 1) * (r + 1) <= n) ++r;
    return (u128)r * r < n ? r + 1 : r;
}

/* run_bsgs_reference (Math.hpp:488-542): plain BSGS over n in [0, m*m), used for p < 1000,
 * p >= 2^62, even p and p | b. */
static int run_bsgs_plain(uint64_t prime, const or_item* item, or_relation* out) {
    if (prime < 2) return 0;
    const uint64_t b_mod = mod_i64(item->b, prime);
    const uint64_t t = neg_mod_i64(item->c, prime);
    if (b_mod == 0) {
        if (t == 0) {
            out->p = prime;
            out->M = (uint64_t)item->N0;
            out->q = 1;
            return 1;
        }
        return 0;
    }
    const uint64_t m = (prime > (uint64_t)item->N1) ? isqrt_ceil_u64((uint64_t)item->N1) : isqrt_ceil_u64(prime);
    if (m == 0) return 0;
    table_t* tab = &tls_table;
    table_reset(tab, m);
    uint64_t power = 1;
    for (uint64_t i = 0; i < m; ++i) {
        table_insert(tab, power, (uint32_t)i);
        power = mulmod(power, b_mod, prime);
    }
    const uint64_t step = or_powmod(b_mod, m, prime);
    const uint64_t step_inv = or_powmod(step, prime - 2, prime);
    uint64_t giant = t;
    for (uint64_t j = 0; j < m; ++j) {
        uint32_t idx;
        if (table_find(tab, giant, &idx)) {
            out->p = prime;
            out->M = idx + j * m;
            out->q = or_multiplicative_order(b_mod, prime);
            return 1;
        }
        giant = mulmod(giant, step_inv, prime);
    }
    return 0;
}

/* run_bsgs (Math.hpp:377-486).  Searches n = n_start + cm*k, k in [0, kcount): B^k == T with
 * B = b^cm, T = -c * b^-n_start.  M is the smallest in-class solution n >= N0. */
int or_run_bsgs(uint64_t prime, const or_item* item, or_relation* out) {
    if (prime < 1000 || prime >= (1ULL << 62) || (prime & 1) == 0) return run_bsgs_plain(prime, item, out);
    const uint64_t b_mod = mod_i64(item->b, prime);
    const uint64_t t = neg_mod_i64(item->c, prime);
    if (b_mod == 0) return run_bsgs_plain(prime, item, out);
    if (t == 0) return 0;

    const uint64_t N0 = item->N0 < 0 ? 0 : (uint64_t)item->N0;
    const uint64_t N1 = (uint64_t)item->N1;
    const uint64_t cm = item->cls_mod == 0 ? 1 : item->cls_mod;
    const uint64_t n_start = N0 + ((item->cls_res % cm) + cm - (N0 % cm)) % cm;
    if (n_start >= N1) return 0;
    uint64_t kcount = (N1 - n_start + cm - 1) / cm;
    if (kcount > prime) kcount = prime;

    mont_t mt;
    mont_init(&mt, prime);
    const uint64_t b_m = mont_to(&mt, b_mod);
    uint64_t B = mont_pow(&mt, b_m, cm);
    uint64_t target = mont_to(&mt, t);
    if (n_start != 0) target = mont_mul(&mt, target, mont_pow(&mt, mont_pow(&mt, b_m, n_start), prime - 2));

    uint64_t stride = 1, parity = 0;
    const uint64_t g = gcd_u64(cm, prime - 1);
    if (g >= 2) {
        if (mont_pow(&mt, target, (prime - 1) / g) != mt.one) return 0;
    } else {
        const int lb = mont_legendre(&mt, B);
        const int lt = mont_legendre(&mt, target);
        if (lb == 1) {
            if (lt == -1) return 0;
        } else {
            stride = 2;
            if (lt == -1) {
                parity = 1;
                target = mont_mul(&mt, target, mont_pow(&mt, B, prime - 2));
            }
            B = mont_mul(&mt, B, B);
        }
    }
    if (kcount <= parity) return 0;
    const uint64_t jcount = (kcount - parity + stride - 1) / stride;
    uint64_t m = isqrt_ceil_u64(jcount);
    if (m == 0) m = 1;

    table_t* tab = &tls_table;
    table_reset(tab, m);
    uint64_t power = mt.one, cycle = 0;
    for (uint64_t i = 0; i < m; ++i) {
        if (i > 0 && power == mt.one) {
            cycle = i;
            break;
        }
        table_insert(tab, power, (uint32_t)i);
        power = mont_mul(&mt, power, B);
    }
    uint64_t j = 0;
    int found = 0;
    uint32_t idx = 0;
    if (cycle != 0) {
        if (table_find(tab, target, &idx)) {
            j = idx;
            found = 1;
        }
    } else {
        const uint64_t step_inv = mont_pow(&mt, power, prime - 2);
        uint64_t giant = target;
        for (uint64_t gs = 0; gs * m < jcount; ++gs) {
            if (table_find(tab, giant, &idx)) {
                j = idx + gs * m;
                found = 1;
                break;
            }
            giant = mont_mul(&mt, giant, step_inv);
        }
    }
    if (!found || j >= jcount) return 0;
    out->p = prime;
    out->M = n_start + (j * stride + parity) * cm;
    out->q = or_multiplicative_order(b_mod, prime);
    return 1;
}

/* ---------------------------------------------------------------- sieves */

static uint32_t* g_base = NULL; /* odd primes 3 .. g_base_limit */
static size_t g_nbase = 0;
static uint64_t g_base_limit = 0;
static pthread_mutex_t g_base_mu = PTHREAD_MUTEX_INITIALIZER;

/* Odd primes up to limit (inclusive), cached: sieve_small_primes (Sieve.hpp:13-39) gives the same
 * list.  The odd numbers are sieved in segments of 2^18 by the odd primes up to sqrt(limit), which
 * come from a plain odd-only sieve (limit < 2^32, so they are below 2^16). */
static void ensure_base_primes(uint64_t limit) {
    pthread_mutex_lock(&g_base_mu);
    if (g_base != NULL && g_base_limit >= limit) {
        pthread_mutex_unlock(&g_base_mu);
        return;
    }
    if (limit < 65536) limit = 65536;
    uint64_t r = 1;
    while ((r + 1) * (r + 1) <= limit) ++r;
    const uint64_t rodd = r / 2 + 1;
    uint8_t* rc = (uint8_t*)calloc(rodd, 1);
    uint32_t* rp = (uint32_t*)malloc(rodd * sizeof(uint32_t));
    size_t nrp = 0;
    for (uint64_t i = 1; i < rodd; ++i) {
        if (rc[i]) continue;
        const uint64_t q = 2 * i + 1;
        rp[nrp++] = (uint32_t)q;
        for (uint64_t j = (q * q) / 2; j < rodd; j += q) rc[j] = 1;
    }
    free(rc);
    size_t cap = 1 << 16, cnt = 0;
    uint32_t* pr = (uint32_t*)malloc(cap * sizeof(uint32_t));
    const uint64_t nodd = limit / 2 + 1; /* index i <-> 2i+1, i = 1 .. */
    const uint64_t seg = 1u << 18;
    uint8_t* comp = (uint8_t*)malloc(seg);
    for (uint64_t s0 = 1; s0 < nodd; s0 += seg) {
        const uint64_t len = nodd - s0 < seg ? nodd - s0 : seg;
        memset(comp, 0, len);
        for (size_t k = 0; k < nrp; ++k) {
            const uint64_t q = rp[k];
            uint64_t j = (q * q) / 2; /* index of q^2 */
            if (j >= s0 + len) break;
            if (j < s0) j += ((s0 - j + q - 1) / q) * q;
            for (; j < s0 + len; j += q) comp[j - s0] = 1;
        }
        for (uint64_t i = 0; i < len; ++i) {
            if (comp[i] || 2 * (s0 + i) + 1 > limit) continue;
            if (cnt == cap) {
                cap *= 2;
                pr = (uint32_t*)realloc(pr, cap * sizeof(uint32_t));
            }
            pr[cnt++] = (uint32_t)(2 * (s0 + i) + 1);
        }
    }
    free(comp);
    free(rp);
    free(g_base);
    g_base = pr;
    g_nbase = cnt;
    g_base_limit = limit;
    pthread_mutex_unlock(&g_base_mu);
}

#define SEG_ODDS (1u << 18)

typedef struct {
    uint32_t q, pos;
} hit_t;
typedef struct {
    hit_t* v;
    size_t n, cap;
} hitvec_t;

static void hv_push(hitvec_t* h, uint32_t q, uint32_t pos) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 64;
        h->v = (hit_t*)realloc(h->v, h->cap * sizeof(hit_t));
    }
    h->v[h->n].q = q;
    h->v[h->n].pos = pos;
    h->n++;
}

typedef void (*emit_fn)(void* ctx, uint64_t v);

/* Odd x in [lo, hi) (lo odd, > 3) not crossed off by any q in primes[0..np): each q crosses off
 * its odd multiples from max(q^2, the first odd multiple >= lo).  A bucket sieve carries the
 * primes larger than a segment.  Calls emit in ascending order. */
static void sieve_odd(uint64_t lo, uint64_t hi, const uint32_t* primes, size_t np, emit_fn emit, void* ctx) {
    if (lo >= hi) return;
    const uint64_t count = (hi - lo + 1) / 2;
    const uint64_t nseg = (count + SEG_ODDS - 1) / SEG_ODDS;
    uint32_t* sq = (uint32_t*)malloc(np * sizeof(uint32_t) + 1);
    uint64_t* snext = (uint64_t*)malloc(np * sizeof(uint64_t) + 1);
    size_t nsmall = 0;
    hitvec_t* buckets = (hitvec_t*)calloc(nseg, sizeof(hitvec_t));
    for (size_t k = 0; k < np; ++k) {
        const uint64_t q = primes[k];
        const uint64_t qq = q * q;
        if (qq >= hi) continue;
        uint64_t start = ((lo + q - 1) / q) * q;
        if ((start & 1) == 0) start += q;
        if (start < qq) start = qq;
        const uint64_t idx = (start - lo) / 2;
        if (idx >= count) continue;
        if (q < SEG_ODDS) {
            sq[nsmall] = (uint32_t)q;
            snext[nsmall] = idx;
            nsmall++;
        } else {
            hv_push(&buckets[idx / SEG_ODDS], (uint32_t)q, (uint32_t)(idx % SEG_ODDS));
        }
    }
    uint8_t* seg = (uint8_t*)malloc(SEG_ODDS);
    for (uint64_t s = 0; s < nseg; ++s) {
        const uint64_t base = s * SEG_ODDS;
        const uint64_t len = (count - base) < SEG_ODDS ? (count - base) : SEG_ODDS;
        memset(seg, 0, len);
        const uint64_t end = base + len;
        for (size_t k = 0; k < nsmall; ++k) {
            const uint64_t q = sq[k];
            uint64_t v = snext[k];
            for (; v < end; v += q) seg[v - base] = 1;
            snext[k] = v;
        }
        hitvec_t* bk = &buckets[s];
        for (size_t i = 0; i < bk->n; ++i) {
            const uint64_t q = bk->v[i].q;
            const uint64_t pos = base + bk->v[i].pos;
            seg[pos - base] = 1;
            const uint64_t nx = pos + q;
            if (nx < count) hv_push(&buckets[nx / SEG_ODDS], (uint32_t)q, (uint32_t)(nx % SEG_ODDS));
        }
        free(bk->v);
        bk->v = NULL;
        bk->n = bk->cap = 0;
        const uint64_t vbase = lo + 2 * base;
        for (uint64_t i = 0; i < len; ++i)
            if (!seg[i]) emit(ctx, vbase + 2 * i);
    }
    free(seg);
    free(buckets);
    free(sq);
    free(snext);
}

typedef struct {
    uint64_t* out;
    size_t cap, n;
} collect_t;

static void emit_collect(void* ctx, uint64_t v) {
    collect_t* c = (collect_t*)ctx;
    if (c->out != NULL && c->n < c->cap) c->out[c->n] = v;
    c->n++;
}

static uint64_t isqrt_floor_u64(uint64_t n) {
    uint64_t r = isqrt_ceil_u64(n);
    return (u128)r * r > n ? r - 1 : r;
}

/* segmented_sieve (Sieve.hpp:110-185): the primes of [lo, hi) in ascending order. */
size_t or_sieve_primes(uint64_t lo, uint64_t hi, uint64_t* out, size_t cap) {
    collect_t c = {out, cap, 0};
    if (lo >= hi) return 0;
    if (lo <= 2 && hi > 2) emit_collect(&c, 2);
    uint64_t olo = lo < 3 ? 3 : (lo | 1);
    if (olo >= hi) return c.n;
    const uint64_t lim = hi > 1 ? isqrt_floor_u64(hi - 1) : 1;
    ensure_base_primes(lim);
    size_t np = 0;
    while (np < g_nbase && (uint64_t)g_base[np] <= lim) ++np;
    sieve_odd(olo, hi, g_base, np, emit_collect, &c);
    return c.n;
}

/* psieve (Gpu.hpp:385-455) output set, for lo > 3. */
size_t or_sieve_survivors(uint64_t lo, uint64_t hi, uint64_t* out, size_t cap) {
    collect_t c = {out, cap, 0};
    ensure_base_primes(65536);
    size_t np = 0;
    while (np < g_nbase && g_base[np] < 65536) ++np; /* 6541 odd primes 3 .. 65521 */
    sieve_odd(lo | 1, hi, g_base, np, emit_collect, &c);
    return c.n;
}

/* ---------------------------------------------------------------- device-stage decisions */

/* gpu_classes (Math.hpp:324-337). */
int or_gpu_classes(const or_item* item, uint64_t* Q, uint32_t* classes, uint32_t* nclasses) {
    const uint64_t cm = item->cls_mod == 0 ? 1 : item->cls_mod;
    *nclasses = 0;
    if (item->N0 < 0 || item->N1 <= item->N0) return 0;
    const uint64_t span = (uint64_t)(item->N1 - item->N0);
    *Q = cm / gcd_u64(cm, 24) * 24;
    if ((cm & 1) != 0 || *Q > 48 || span / *Q < 4096) return 0;
    uint32_t n = 0;
    for (uint64_t r = 0; r < *Q; ++r)
        if (r % cm == item->cls_res % cm) classes[n++] = (uint32_t)r;
    *nclasses = n;
    return n > 0 && n <= 24;
}

static int fermat2(const mont_t* mt) { /* 2^(p-1) == 1 (mod p) */
    return mont_pow(mt, mont_to(mt, 2), mt->n - 1) == mt->one;
}

/* prep_core (Gpu.hpp:231-308): residue filters, base-2 Fermat test (need_prime), g-th power test. */
int or_prep(uint64_t p, const or_item* item, int need_prime, or_entry* out) {
    if (p < 1000 || (p & 1) == 0) return 2;
    const uint64_t b_mod = (uint64_t)item->b % p;
    const uint64_t t = neg_mod_i64(item->c, p);
    if (b_mod == 0) return 2;
    if (t == 0) return 0;
    const uint64_t N0 = item->N0 < 0 ? 0 : (uint64_t)item->N0;
    const uint64_t N1 = (uint64_t)item->N1;
    const uint64_t cm = item->cls_mod == 0 ? 1 : item->cls_mod;
    const uint64_t cres = item->cls_res;
    const uint64_t n_start = N0 + ((cres % cm) + cm - (N0 % cm)) % cm;
    if (n_start >= N1) return 0;
    uint64_t kcount = (N1 - n_start + cm - 1) / cm;
    if (kcount > p) kcount = p;
    uint32_t stride = 1, parity = 0;
    const uint64_t g = gcd_u64(cm, p - 1);
    const int jb = or_jacobi(b_mod, p);
    const int jt = or_jacobi(t, p);
    const int lT = jt * ((n_start & 1) ? jb : 1);
    if (g == 2) {
        if (lT != 1) return 0;
    } else if (g < 2) {
        if (jb == 1) {
            if (lT != 1) return 0;
        } else {
            stride = 2;
            if (lT != 1) parity = 1;
        }
    }
    mont_t mt;
    mont_init(&mt, p);
    if (need_prime && !fermat2(&mt)) return 0;
    if (g > 2) {
        const uint64_t e = (p - 1) / g;
        const uint64_t tm = mont_to(&mt, t);
        const uint64_t G0 = mont_pow(&mt, mont_to(&mt, b_mod), n_start);
        if (mont_pow(&mt, tm, e) != mont_pow(&mt, G0, e)) return 0;
    }
    if (kcount <= parity) return 0;
    const uint64_t jcount = (kcount - parity + stride - 1) / stride;
    uint64_t m = isqrt_ceil_u64(jcount);
    if (m == 0) m = 1;
    out->jcount = jcount;
    out->m = m;
    out->stride = stride;
    out->parity = parity;
    return 1;
}

/* prep_q_core (Gpu.hpp:478-589): per-prime class split by qp = lcm(cm, gcd(Qmax, p-1)). */
int or_prep_q(uint64_t p, const or_item* item, uint64_t Qmax, int need_prime, or_entry_q* out) {
    if (p < 1000 || (p & 1) == 0) return 2;
    const uint64_t b_mod = (uint64_t)item->b % p;
    const uint64_t t = neg_mod_i64(item->c, p);
    if (b_mod == 0) return 2;
    if (t == 0) return 0;
    const uint64_t N0 = item->N0 < 0 ? 0 : (uint64_t)item->N0;
    const uint64_t N1 = (uint64_t)item->N1;
    const uint64_t cm = item->cls_mod == 0 ? 1 : item->cls_mod;
    const uint64_t span = N1 - N0;
    const uint64_t g = gcd_u64(Qmax, p - 1);
    const uint32_t qp = (uint32_t)(g / gcd_u64(g, cm) * cm);
    const uint32_t ncls = qp / (uint32_t)cm;
    const uint32_t n0q = (uint32_t)(N0 % qp);
    const uint32_t c0 = (uint32_t)(item->cls_res % cm);
    uint32_t dj[48];
    for (uint32_t j = 0; j < ncls; ++j) dj[j] = (c0 + j * (uint32_t)cm + qp - n0q) % qp;
    const int jb = or_jacobi(b_mod, p);
    const int jt = or_jacobi(t, p);
    const uint32_t n0odd = (uint32_t)(N0 & 1);
    uint32_t mask = 0;
    for (uint32_t j = 0; j < ncls; ++j) {
        if ((uint64_t)dj[j] >= span) continue;
        const int l = jt * (((n0odd ^ (dj[j] & 1)) != 0) ? jb : 1);
        if (l == 1) mask |= 1u << j;
    }
    if (mask == 0) return 0;
    mont_t mt;
    mont_init(&mt, p);
    if (need_prime && !fermat2(&mt)) return 0;
    if (g > 2) {
        const uint64_t e = (p - 1) / g;
        const uint64_t u = mont_pow(&mt, mont_to(&mt, t), e);
        const uint64_t w = mont_pow(&mt, mont_to(&mt, b_mod), e);
        const uint32_t n0g = (uint32_t)(N0 % g);
        uint32_t keep = 0;
        uint64_t wk = mt.one;
        for (uint32_t k = 0; k < (uint32_t)g; ++k) {
            if (wk == u)
                for (uint32_t j = 0; j < ncls; ++j)
                    if ((n0g + dj[j]) % (uint32_t)g == k) keep |= 1u << j;
            wk = mont_mul(&mt, wk, w);
        }
        mask &= keep;
        if (mask == 0) return 0;
    }
    uint32_t s = 0;
    for (uint32_t j = 0; j < ncls; ++j) s += (mask >> j) & 1;
    /* baby/giant split (Gpu.hpp:553-573): Gc = 64*cr giant steps per class, m <= 1024 */
    const uint64_t K = (span + qp - 1) / qp;
    uint64_t crmin = (K + 64 * 1023 - 1) / (64 * 1023);
    if (crmin < 2) crmin = 2;
    const uint64_t crc = (uint64_t)sqrt((double)K / (64.0 * 64.0 * (double)s));
    uint64_t best = ~0ULL, m = 1024, Gc = 64 * crmin;
    for (uint64_t cr = (crc > 1 ? crc - 1 : 1); cr <= crc + 2; ++cr) {
        const uint64_t cx = cr < crmin ? crmin : cr;
        const uint64_t gc = 64 * cx;
        const uint64_t mm = (K + gc - 3) / (gc - 1) + 1;
        if (mm > 1024) continue;
        const uint64_t rounds = (mm + 63) / 64 + (uint64_t)s * cx;
        if (rounds < best) {
            best = rounds;
            m = mm;
            Gc = gc;
        }
    }
    if (best == ~0ULL) {
        for (uint64_t cx = crmin;; ++cx) {
            const uint64_t gc = 64 * cx;
            const uint64_t mm = (K + gc - 3) / (gc - 1) + 1;
            if (mm <= 1024) {
                m = mm;
                Gc = gc;
                break;
            }
        }
    }
    out->qp = qp;
    out->ncls = ncls;
    out->mask = mask;
    out->s = s;
    out->m = m;
    out->Gc = Gc;
    return 1;
}

/* ---------------------------------------------------------------- CPU branch, threaded */

typedef struct {
    const or_item* item;
    uint64_t lo, hi, chunk;
    uint64_t nchunks;
    uint64_t next; /* atomic */
    or_relation* rels;
    size_t nrel, cap; /* guarded by mu */
    uint64_t primes;
    pthread_mutex_t mu;
} prange_t;

typedef struct {
    const or_item* item;
    or_relation* v;
    size_t n, cap;
    uint64_t primes;
} bsgs_ctx_t;

static void emit_bsgs(void* ctx, uint64_t p) {
    bsgs_ctx_t* c = (bsgs_ctx_t*)ctx;
    c->primes++;
    or_relation r;
    if (or_run_bsgs(p, c->item, &r)) {
        if (c->n == c->cap) {
            c->cap = c->cap ? 2 * c->cap : 64;
            c->v = (or_relation*)realloc(c->v, c->cap * sizeof(or_relation));
        }
        c->v[c->n++] = r;
    }
}

static void* prange_worker(void* arg) {
    prange_t* pr = (prange_t*)arg;
    for (;;) {
        const uint64_t k = __atomic_fetch_add(&pr->next, 1, __ATOMIC_RELAXED);
        if (k >= pr->nchunks) break;
        const uint64_t clo = pr->lo + k * pr->chunk;
        const uint64_t chi = (pr->hi - clo) < pr->chunk ? pr->hi : clo + pr->chunk;
        bsgs_ctx_t c = {pr->item, NULL, 0, 0, 0};
        if (clo <= 2 && chi > 2) emit_bsgs(&c, 2);
        const uint64_t olo = clo < 3 ? 3 : (clo | 1);
        size_t np = 0;
        const uint64_t lim = isqrt_floor_u64(chi - 1);
        while (np < g_nbase && (uint64_t)g_base[np] <= lim) ++np;
        if (olo < chi) sieve_odd(olo, chi, g_base, np, emit_bsgs, &c);
        pthread_mutex_lock(&pr->mu);
        for (size_t i = 0; i < c.n; ++i) {
            if (pr->nrel < pr->cap) pr->rels[pr->nrel] = c.v[i];
            pr->nrel++;
        }
        pr->primes += c.primes;
        pthread_mutex_unlock(&pr->mu);
        free(c.v);
    }
    return NULL;
}

static int cmp_rel(const void* a, const void* b) {
    const or_relation* x = (const or_relation*)a;
    const or_relation* y = (const or_relation*)b;
    return x->p < y->p ? -1 : (x->p > y->p ? 1 : 0);
}

size_t or_process_range(const or_item* item, uint64_t lo, uint64_t hi, uint64_t chunk_ints, int nthreads,
                        or_relation* out, size_t cap, uint64_t* primes_seen) {
    if (lo >= hi) return 0;
    if (chunk_ints == 0) chunk_ints = 1ULL << 26;
    if (nthreads < 1) nthreads = 1;
    ensure_base_primes(isqrt_floor_u64(hi - 1));
    prange_t pr;
    memset(&pr, 0, sizeof(pr));
    pr.item = item;
    pr.lo = lo;
    pr.hi = hi;
    pr.chunk = chunk_ints;
    pr.nchunks = (hi - lo + chunk_ints - 1) / chunk_ints;
    pr.rels = out;
    pr.cap = cap;
    pthread_mutex_init(&pr.mu, NULL);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, prange_worker, &pr);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&pr.mu);
    if (primes_seen) *primes_seen = pr.primes;
    if (pr.nrel > cap) return (size_t)-1;
    qsort(out, pr.nrel, sizeof(or_relation), cmp_rel);
    return pr.nrel;
}

/* ---------------------------------------------------------------- results.json */

typedef struct {
    char* s;
    size_t n, cap;
} sbuf_t;

static void sb_put(sbuf_t* b, const char* s, size_t len) {
    if (b->n + len + 1 > b->cap) {
        size_t nc = b->cap ? b->cap : 4096;
        while (b->n + len + 1 > nc) nc *= 2;
        b->s = (char*)realloc(b->s, nc);
        b->cap = nc;
    }
    memcpy(b->s + b->n, s, len);
    b->n += len;
    b->s[b->n] = 0;
}

static void sb_str(sbuf_t* b, const char* s) { sb_put(b, s, strlen(s)); }

static void sb_i128(sbuf_t* b, i128 v) { /* json_int (Results.hpp:57-63) for values in int64 */
    char tmp[64];
    if (v > (i128)INT64_MAX || v < (i128)INT64_MIN) {
        /* outside int64: a JSON string of the decimal value */
        char d[48];
        int k = 0;
        const int neg = v < 0;
        u128 a = neg ? (u128)(-v) : (u128)v;
        do {
            d[k++] = (char)('0' + (int)(a % 10));
            a /= 10;
        } while (a);
        int o = 0;
        tmp[o++] = '"';
        if (neg) tmp[o++] = '-';
        while (k) tmp[o++] = d[--k];
        tmp[o++] = '"';
        tmp[o] = 0;
    } else {
        snprintf(tmp, sizeof(tmp), "%lld", (long long)v);
    }
    sb_str(b, tmp);
}

static i128 ceil_div_i(i128 a, i128 b) { return a >= 0 ? (a + b - 1) / b : a / b; }
static i128 floor_div_i(i128 a, i128 b) { return a >= 0 ? a / b : (a - b + 1) / b; }

static int in_class(const or_item* item, i128 n) { /* Types.hpp:28-31 */
    if (item->cls_mod <= 1) return 1;
    return (uint64_t)(n % (i128)item->cls_mod) == item->cls_res;
}

size_t or_results_json(const or_item* item, const or_relation* rels, size_t nrel, const char* timestamp,
                       char* buf, size_t cap) {
    const i128 N0 = item->N0, N1 = item->N1;
    const size_t count = (size_t)(N1 - N0);
    uint32_t* cnt = (uint32_t*)calloc(count + 1, sizeof(uint32_t));
    /* apply_relations (Worker.hpp:527-559), two passes: count, then fill (CSR) */
    for (int pass = 0; pass < 2; ++pass) {
        static __thread uint64_t* fac;
        static __thread size_t* start;
        if (pass == 1) {
            start = (size_t*)malloc((count + 1) * sizeof(size_t));
            size_t acc = 0;
            for (size_t i = 0; i < count; ++i) {
                start[i] = acc;
                acc += cnt[i];
                cnt[i] = 0;
            }
            start[count] = acc;
            fac = (uint64_t*)malloc((acc + 1) * sizeof(uint64_t));
        }
        for (size_t r = 0; r < nrel; ++r) {
            const i128 q = (i128)rels[r].q, M = (i128)rels[r].M, p = (i128)rels[r].p;
            const i128 x0 = ceil_div_i(N0 - M, q), x1 = floor_div_i(N1 - 1 - M, q);
            for (i128 x = x0; x <= x1; ++x) {
                const i128 n = q * x + M;
                const i128 index = n - N0;
                if (index < 0 || index >= (i128)count) continue;
                /* b^n + c == p: the number itself is the prime, not a factored candidate */
                i128 value = 1;
                int fits = 1;
                for (i128 e = 0; e < n && fits; ++e) {
                    value *= item->b;
                    if (value > p * 2) fits = 0;
                }
                if (fits && value + item->c == p) continue;
                if (pass == 0)
                    cnt[(size_t)index]++;
                else
                    fac[start[(size_t)index] + cnt[(size_t)index]++] = rels[r].p;
            }
        }
        if (pass == 1) {
            sbuf_t b = {NULL, 0, 0};
            sb_str(&b, "{\"Timestamp\":\"");
            sb_str(&b, timestamp);
            sb_str(&b, "\",\"b\":");
            sb_i128(&b, item->b);
            sb_str(&b, ",\"c\":");
            sb_i128(&b, item->c);
            sb_str(&b, ",\"P0\":");
            sb_i128(&b, (i128)item->P0);
            sb_str(&b, ",\"P1\":");
            sb_i128(&b, (i128)item->P1);
            sb_str(&b, ",\"N0\":");
            sb_i128(&b, N0);
            sb_str(&b, ",\"N1\":");
            sb_i128(&b, N1);
            sb_str(&b, ",\"prp\":[");
            int first = 1;
            for (size_t i = 0; i < count; ++i) {
                const i128 n = N0 + (i128)i;
                if (!in_class(item, n) || cnt[i] != 0) continue;
                if (!first) sb_str(&b, ",");
                first = 0;
                sb_i128(&b, n);
            }
            sb_str(&b, "],\"factored\":[");
            first = 1;
            for (size_t i = 0; i < count; ++i) {
                const i128 n = N0 + (i128)i;
                if (!in_class(item, n) || cnt[i] == 0) continue;
                if (!first) sb_str(&b, ",");
                first = 0;
                sb_str(&b, "{\"n\":");
                sb_i128(&b, n);
                sb_str(&b, ",\"factors\":[");
                for (uint32_t k = 0; k < cnt[i]; ++k) {
                    if (k) sb_str(&b, ",");
                    sb_i128(&b, (i128)fac[start[i] + k]); /* json_uint: factors < 2^63 */
                }
                sb_str(&b, "]}");
            }
            sb_str(&b, "]}");
            const size_t len = b.n;
            if (buf != NULL && cap > len) memcpy(buf, b.s, len + 1);
            free(b.s);
            free(fac);
            free(start);
            free(cnt);
            return len;
        }
    }
    free(cnt);
    return 0;
}
