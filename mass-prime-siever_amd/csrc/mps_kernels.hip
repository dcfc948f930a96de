// HIP kernels of the MI355X sieve worker (gfx950, wave64).
//
// Stages, one pass = one range of odd numbers [lo, lo + 2*nodd):
//   k_sieve   segmented odd sieve, one 1024-thread workgroup per 2^17-odd LDS bitmap: the primes
//             3..61 as register combs, 67..4093 a wavefront per prime, 4099..65521 a lane per prime
//             (LDS atomic ORs); every odd multiple is crossed off and a segment restores the
//             sieving primes it holds; survivors are compacted by one block reservation.
//   k_filter  per survivor: quadratic-character filter (plain and order mode) or per-prime class
//             split with a quadratic-residue mask (class mode); block compaction of what is left.
//   k_setup   per candidate: base-2 Fermat test, g-th power residue test (class mode), BSGS sizes
//             and the entry (powers tables, giant starts), compacted by block ballot.
//   k_setup_order  order mode: per candidate g, b^e, t^e; per (candidate, prime of g) pair a
//             Pohlig-Hellman step; per candidate the CRT class; per kept candidate the entry.
//   k_bsgs    one workgroup (1, 2 or 4 wavefronts) per entry: baby steps t*B^i into an LDS
//             fingerprint table, giant steps G_j*(B^m)^g looked up, every match verified exactly;
//             the smallest solution goes to the records.
//   k_prime_records  12-base Miller-Rabin over the pass's records.
// Numbers the device cannot decide (p < 1000, p | b, a short cycle) go to the records for the host.
#include <algorithm>
#include <type_traits>

#include "mps_device.hpp"
#include "mps_kernels.h"

namespace mps {

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Wave-aggregated append: every lane of the wavefront must call it.  Returns the slot of a keeping
// lane (ordered by lane) in the array counted by *counter.
__device__ __forceinline__ uint32_t wave_append(bool keep, uint32_t* counter) {
    const uint64_t mask = __ballot(keep);
    if (mask == 0) return 0;
    const uint32_t lane = lane_id();
    const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __builtin_amdgcn_readlane(base, leader);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    return base + rank;
}

// Block-aggregated append: every thread of the workgroup must call it (it synchronises the
// block).  One atomic per call and workgroup; sh needs NW + 1 words of LDS.
template <int NW>
__device__ __forceinline__ uint32_t block_append(bool keep, uint32_t* counter, uint32_t* sh) {
    const uint64_t mask = __ballot(keep);
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < NW; ++i) {
            const uint32_t c = sh[i];
            sh[i] = tot;
            tot += c;
        }
        sh[NW] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    const uint32_t slot = sh[NW] + sh[w] + rank;
    __syncthreads();
    return slot;
}

// Block-wide reservation of c slots per thread (c may be 0): every thread of the workgroup must
// call it.  One atomic per call and workgroup; returns this thread's first slot.  sh needs NW + 1
// words of LDS.
template <int NW>
__device__ __forceinline__ uint32_t block_reserve(uint32_t c, uint32_t* counter, uint32_t* sh) {
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t x = c;  // inclusive scan over the wavefront
    for (int dd = 1; dd < 64; dd <<= 1) {
        const uint32_t y = __shfl_up(x, dd, 64);
        if (lane >= (uint32_t)dd) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < NW; ++i) {
            const uint32_t t = sh[i];
            sh[i] = tot;
            tot += t;
        }
        sh[NW] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t base = sh[NW] + sh[w] + x - c;
    __syncthreads();
    return base;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Clears the stage counters (a kernel, not a memset, so that it can sit in a captured graph);
// n_surv starts at n_surv0 (the filter path's number count).
__global__ void k_reset(Counters* c, uint32_t n_surv0) {
    uint32_t* w = reinterpret_cast<uint32_t*>(c);
    for (uint32_t i = threadIdx.x; i < sizeof(Counters) / 4; i += blockDim.x) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) c->n_surv = n_surv0;
}

// ------------------------------------------------------------------------------------- sieve
//
// One workgroup of kSieveThreads threads sieves one segment of kSegOdds odd numbers v0, v0 + 2, ...
// in an LDS bitmap (bit i set: v0 + 2i has a sieving prime factor).  Thread t owns bits
// [128t, 128t + 128): four words, one 128-bit LDS access.
//   comb primes (q < kCombPrimeBound): the marks of a word are the prime's comb (bits 0, q, 2q, ...
//     of a 64-bit word) shifted to the word's first multiple; a thread builds its four words in
//     registers over all comb primes and stores them once, without atomics;
//   wave primes (q < kWavePrimeBound): a wavefront per prime, its lanes on consecutive multiples;
//   lane primes (the rest, up to 65521): a lane per prime.
// Every prime crosses off all of its odd multiples in the segment, itself included.  A multiple
// k*q with 1 < k < q has a prime factor below q, which crosses it off as well, so the survivors are
// those of crossing off from q^2 once the sieving primes themselves are restored (only segments
// that start below 65536 hold any).  Survivors, the odd numbers with no sieving prime factor other
// than themselves, are written as indices from the pass start, one block reservation per segment.

// Index j of the first odd multiple v0 + 2j of the odd prime q, from r = v0 mod q: 2j == -r (mod q),
// so j = t / 2 for t = (q - r) mod q when t is even, else (t + q) / 2.
__device__ __forceinline__ uint32_t first_odd_multiple(uint32_t r, uint32_t q) {
    const uint32_t t = r == 0 ? 0u : q - r;
    return (t & 1u) ? (t + q) >> 1 : t >> 1;
}

__device__ __forceinline__ uint32_t seg_first(uint64_t v0, const SievePrime& sp) {
    return first_odd_multiple(mod_small_f(v0, sp.q, sp.r32, sp.inv), sp.q);
}

__global__ __launch_bounds__(kSieveThreads) void k_sieve(const PassDesc* __restrict__ d, SieveTable tab, PassBuffers pb) {
    constexpr uint32_t kWords = kSegOdds / 32, kOwn = 4;  // words per thread
    static_assert(kWords == kOwn * kSieveThreads, "a thread owns four words of the segment");
    __shared__ uint4 bits4[kWords / 4];
    __shared__ SievePrime c_sp[kCombPrimes];
    __shared__ uint32_t c_first[kCombPrimes], c_step[kCombPrimes];
    __shared__ unsigned long long c_comb[kCombPrimes];
    __shared__ uint32_t sh[kSieveThreads / 64 + 1];
    uint32_t* bits = reinterpret_cast<uint32_t*>(bits4);
    const uint64_t nodd = d->nodd;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint64_t seg0 = (uint64_t)blockIdx.x * kSegOdds;
    if (seg0 >= nodd) return;  // the whole workgroup (the grid is sized for the largest pass)
    const uint32_t len = nodd - seg0 < kSegOdds ? (uint32_t)(nodd - seg0) : kSegOdds;
    const uint64_t v0 = d->lo + 2 * seg0;

    // the comb primes' first multiples, 32 mod q (the shift from one word to the next) and combs
    if (tid < tab.n_comb) {
        const SievePrime sp = tab.sp[tid];
        c_sp[tid] = sp;
        c_first[tid] = seg_first(v0, sp);
        c_step[tid] = 32u % sp.q;
        unsigned long long comb = 0;
        for (uint32_t b = 0; b < 64; b += sp.q) comb |= 1ull << b;
        c_comb[tid] = sp.q < 32 ? comb : 1ull;
    }
    __syncthreads();
    {
        const uint32_t x = kOwn * 32 * tid;  // the thread's first bit
        uint32_t w[kOwn] = {0u, 0u, 0u, 0u};
        for (uint32_t k = 0; k < tab.n_comb; ++k) {
            const uint32_t q = c_sp[k].q, step = c_step[k];
            const unsigned long long comb = c_comb[k];
            // offset of the first multiple at or after bit x: (first - x) mod q
            uint32_t o = c_first[k] + q - mod16_f(x, q, c_sp[k].inv);
            o = o >= q ? o - q : o;
#pragma unroll
            for (uint32_t i = 0; i < kOwn; ++i) {
                w[i] |= o < 64 ? (uint32_t)(comb << o) : 0u;
                o = o >= step ? o - step : o + q - step;  // offset in the next word
            }
        }
        bits4[tid] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    __syncthreads();
    // wave primes: wavefront `wave` takes every sixteenth one
    for (uint32_t k = tab.n_comb + wave; k < tab.n_wave; k += kSieveThreads / 64) {
        const SievePrime sp = tab.sp[k];
        const uint32_t q = sp.q;
        for (uint32_t i = seg_first(v0, sp) + lane * q; i < len; i += 64 * q) atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    // lane primes
    for (uint32_t k = tab.n_wave + tid; k < tab.n; k += kSieveThreads) {
        const SievePrime sp = tab.sp[k];
        for (uint32_t i = seg_first(v0, sp); i < len; i += sp.q) atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    __syncthreads();
    if (v0 < 65536) {  // the sieving primes inside the segment are primes: restore them
        for (uint32_t k = tid; k < tab.n; k += kSieveThreads) {
            const uint32_t q = tab.sp[k].q;
            if (q >= v0 && (q - v0) / 2 < len) {
                const uint32_t i = (uint32_t)((q - v0) / 2);
                atomicAnd(&bits[i >> 5], ~(1u << (i & 31)));
            }
        }
        __syncthreads();
    }
    // survivors: the clear bits below len, in bit order per thread, after one block reservation
    const uint4 m = bits4[tid];
    uint32_t cw[kOwn] = {~m.x, ~m.y, ~m.z, ~m.w};
    const uint32_t x = kOwn * 32 * tid;
    uint32_t c = 0;
#pragma unroll
    for (uint32_t i = 0; i < kOwn; ++i) {
        const uint32_t b0 = x + 32 * i;  // bits at or past len are not in the segment
        if (b0 >= len) cw[i] = 0;
        else if (len - b0 < 32) cw[i] &= (1u << (len - b0)) - 1u;
        c += __popc(cw[i]);
    }
    uint32_t pos = block_reserve<kSieveThreads / 64>(c, &pb.cnt->n_surv, sh);
    const uint32_t idx0 = (uint32_t)seg0 + x;
#pragma unroll
    for (uint32_t i = 0; i < kOwn; ++i) {
        for (uint32_t v = cw[i]; v != 0; v &= v - 1u, ++pos)
            if (pos < pb.surv_cap) pb.surv[pos] = idx0 + 32 * i + (uint32_t)__builtin_ctz(v);
    }
}

// ------------------------------------------------------------------------------------- filter

__device__ __forceinline__ uint64_t mod_signed(int64_t x, uint64_t p) {  // x mod p in [0, p)
    if (x >= 0) {
        const uint64_t u = (uint64_t)x;
        return u < p ? u : u % p;
    }
    const uint64_t a = 0ull - (uint64_t)x;  // |x|, 2^63 for INT64_MIN
    const uint64_t r = a < p ? a : a % p;
    return r == 0 ? 0 : p - r;
}

// (-x) mod p in [0, p): t = -c of the item, without negating c in 64 bits (-INT64_MIN is 2^63).
__device__ __forceinline__ uint64_t neg_mod(int64_t x, uint64_t p) {
    if (x <= 0) {
        const uint64_t u = 0ull - (uint64_t)x;
        return u < p ? u : u % p;
    }
    const uint64_t u = (uint64_t)x;
    const uint64_t r = u < p ? u : u % p;
    return r == 0 ? 0 : p - r;
}

// (p - 1) mod d for the class modulus: 32-bit path when d < 2^16.
__device__ __forceinline__ uint64_t pm1_mod(uint64_t p, uint64_t d, uint32_t r32, double dinv) {
    if (d < 65536u) return mod_small_f(p - 1, (uint32_t)d, r32, dinv);
    return (p - 1) % d;
}

// (x / p) for x = b or -c of the item (x given as well, for the general path).
__device__ __forceinline__ int jacobi_item(const JacobiPre& J, int64_t x, uint64_t p) {
    if (J.a == 0) return jacobi_signed(x, p);
    int s = 1;
    const uint32_t p8 = (uint32_t)p & 7u;
    if (J.neg && (p8 & 3u) == 3u) s = -s;                    // (-1/p)
    if (J.e_odd && (p8 == 3u || p8 == 5u)) s = -s;           // (2/p)
    if (J.a == 1) return s;
    if ((J.a & 3u) == 3u && (p8 & 3u) == 3u) s = -s;         // reciprocity
    return s * jacobi32(mod_small_f(p, J.a, J.r32, J.dinv), J.a);
}

// (x/p) for x = -c with a table of (r/a), r < a, in LDS (k_filter builds it once per workgroup for
// the odd part a of |c| up to kJacTab): the same signs as jacobi_item, then one lookup of p mod a
// in place of the Jacobi loop.
constexpr uint32_t kJacTab = 4096;
__device__ __forceinline__ int jacobi_item_tab(const JacobiPre& J, int64_t x, uint64_t p, const int8_t* tab) {
    if (tab == nullptr) return jacobi_item(J, x, p);
    int s = 1;
    const uint32_t p8 = (uint32_t)p & 7u;
    if (J.neg && (p8 & 3u) == 3u) s = -s;
    if (J.e_odd && (p8 == 3u || p8 == 5u)) s = -s;
    if ((J.a & 3u) == 3u && (p8 & 3u) == 3u) s = -s;
    return s * tab[mod_small_f(p, J.a, J.r32, J.dinv)];
}

__device__ __forceinline__ uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

__device__ __forceinline__ void put_record(const PassBuffers& pb, uint32_t slot, uint64_t num, uint64_t M) {
    if (slot < pb.out_cap) {
        pb.out[2 * (uint64_t)slot] = num;
        pb.out[2 * (uint64_t)slot + 1] = M;
    }
}

// The residue decisions of one number p for the item (prep_core / prep_q_core, Gpu.hpp:231-263,
// :478-509).  Numbers the device leaves to the host (p < 1000, p even, p | b) set `host`; a number
// that cannot divide any b^n + c (t = -c == 0 mod p, no in-class exponent) keeps nothing.  Both
// decisions rest on the quadratic characters: b^n == t (mod p) needs (t/p) == (b/p)^n.

// Plain and order mode: info bit 0 = stride 2 (only exponents n_start + 2k + parity can solve), bit
// 1 = the odd k (parity 1).  gm = gcd(cm, p - 1).  Where every in-class exponent has n_start's
// parity (cm even) or (b/p) == 1, b^n has one character over the class, which (t/p) must match:
// such a p is kept only when the character of t * b^-n_start is +1.  This holds whatever gm is.
// With gm > 2 the setup's power test decides as well; for an even cm (so an even gm) that test
// implies the character, and order mode's 2-power step tests it, so the filter only drops earlier
// what the setup would drop.  Plain mode with an odd cm and gm > 2 keeps every p, as the
// reference's prep_core does (its gm-th power test is then of odd degree).  With (b/p) == -1 and
// an odd cm the characters fix the parity of k instead (gm == 1; gm > 2 leaves it to the setup).
__device__ __forceinline__ void filter_plain(uint64_t p, const ItemParams& ip, const int8_t* jtab, bool& keep,
                                             bool& host, uint32_t& info) {
    if (p < 1000 || (p & 1) == 0 || mod_signed(ip.b, p) == 0) {
        host = true;
        return;
    }
    if (mod_signed(ip.c, p) == 0 || ip.kcount == 0) return;
    const uint64_t pm = pm1_mod(p, ip.cm, ip.cm_r32, ip.cm_dinv);
    const uint64_t gm = ip.cm < (1ull << 32) ? gcd32((uint32_t)ip.cm, (uint32_t)pm) : gcd64(ip.cm, pm);
    const int sb = jacobi_item(ip.jb, ip.b, p), st = jacobi_item_tab(ip.jt, (int64_t)(0ull - (uint64_t)ip.c), p, jtab);
    // the character of t * b^-n_start: +1 is needed when k = 0 is possible, i.e. for even k
    const int s0 = (ip.n_start & 1) ? st * sb : st;
    const bool even_cm = (ip.cm & 1) == 0;
    if (even_cm || sb == 1) {
        // one character over the class
        keep = (gm > 2 && !even_cm && ip.qmode != kModeOrder) || s0 == 1;
    } else if (gm > 2) {
        keep = true;
    } else {
        keep = true;  // (b/p) == -1, cm odd: k even needs s0 == 1, k odd needs s0 == -1
        info = 1u | (s0 == 1 ? 0u : 2u);
    }
}

// Class mode.  With Q = lcm(cm, 24), g = gcd(Q, p - 1) and qp = lcm(cm, g), the in-class exponents
// n_start + i*cm (i < kcount) fall into ncls = qp / cm classes by i mod ncls: class k holds
// n_start + k*cm + l*qp (l >= 0), all of one parity since qp is even.  Bit k of a class mask stands
// for class k; its first exponent lies k*cm past n_start.
struct ExpClasses {
    uint32_t g, qp, ncls;
};

__device__ __forceinline__ ExpClasses exp_classes(uint64_t p, const ItemParams& ip) {
    ExpClasses ec;
    const uint32_t Q = (uint32_t)ip.Qmax, cm = (uint32_t)ip.cm;
    ec.g = gcd32(Q, mod_small_f(p - 1, Q, ip.Q_r32, ip.Q_dinv));
    ec.qp = ec.g / gcd32(ec.g, cm) * cm;
    ec.ncls = ec.qp / cm;
    return ec;
}

// info = the classes whose parity the characters allow (and that hold an exponent below N1).
__device__ __forceinline__ void filter_class(uint64_t p, const ItemParams& ip, const int8_t* jtab, bool& keep,
                                             bool& host, uint32_t& info) {
    if (p < 1000 || (p & 1) == 0 || mod_signed(ip.b, p) == 0) {
        host = true;
        return;
    }
    if (mod_signed(ip.c, p) == 0) return;
    const ExpClasses ec = exp_classes(p, ip);
    const int sb = jacobi_item(ip.jb, ip.b, p), st = jacobi_item_tab(ip.jt, (int64_t)(0ull - (uint64_t)ip.c), p, jtab);
    const uint32_t nk = ip.kcount < ec.ncls ? (uint32_t)ip.kcount : ec.ncls;
    const uint32_t odd0 = (uint32_t)(ip.n_start & 1), odd_step = (uint32_t)(ip.cm & 1);
    uint32_t mask = 0;
    for (uint32_t k = 0; k < nk; ++k) {
        const bool odd = (odd0 ^ (k & odd_step)) != 0;  // parity of n_start + k*cm
        if (st == (odd ? sb : 1)) mask |= 1u << k;
    }
    keep = mask != 0;
    info = mask;
}

__global__ __launch_bounds__(256) void k_filter(const PassDesc* __restrict__ d, PassBuffers pb) {
    const uint64_t lo = d->lo;
    const ItemParams& ip = d->ip;  // read in place: uniform fields load as scalars, tables by index
    __shared__ uint32_t sh[5];
    // (r/a) for the odd part a of |c|, r < a, when a is small enough
    __shared__ int8_t s_jt[kJacTab];
    const uint32_t ja = ip.jt.a;
    const int8_t* jtab = nullptr;
    if (ja > 1 && ja <= kJacTab) {  // uniform
        for (uint32_t r = threadIdx.x; r < ja; r += 256) s_jt[r] = (int8_t)jacobi32(r, ja);
        __syncthreads();
        jtab = s_jt;
    }
    const uint32_t n = min(pb.cnt->n_surv, pb.surv_cap);
    const uint32_t lane = lane_id();
    uint32_t nflag2 = 0;
    // kFilterPer survivors per thread between two reservations of candidate slots: one atomic
    // on the candidate counter per 256 * kFilterPer survivors (same-address atomics serialise)
    constexpr uint32_t kFilterPer = 4;
    for (uint32_t i0 = blockIdx.x * 256 * kFilterPer; i0 < n; i0 += gridDim.x * 256 * kFilterPer) {
        bool kept[kFilterPer];
        uint64_t val[kFilterPer];
        uint32_t c = 0;
#pragma unroll
        for (uint32_t k = 0; k < kFilterPer; ++k) {
            const uint32_t i = i0 + k * 256 + threadIdx.x;
            bool keep = false, host = false;
            uint32_t info = 0, idx = 0;
            uint64_t p = 0;
            if (i < n) {
                idx = pb.vals ? i : pb.surv[i];
                p = pb.vals ? pb.vals[i] : lo + 2 * (uint64_t)idx;
                if (ip.qmode == kModeClass)
                    filter_class(p, ip, jtab, keep, host, info);
                else
                    filter_plain(p, ip, jtab, keep, host, info);
            }
            kept[k] = keep;
            val[k] = (uint64_t)idx | ((uint64_t)info << 32);
            c += keep ? 1u : 0u;
            const uint32_t oslot = wave_append(host, &pb.cnt->n_out);
            if (host) put_record(pb, oslot, pb.out_index ? idx : p, kHostDecides);
            nflag2 += host ? 1u : 0u;
        }
        uint32_t slot = block_reserve<4>(c, &pb.cnt->n_cand, sh);
#pragma unroll
        for (uint32_t k = 0; k < kFilterPer; ++k) {
            if (kept[k]) {
                if (slot < pb.cand_cap) pb.cand[slot] = val[k];
                ++slot;
            }
        }
    }
    nflag2 = (uint32_t)wave_sum64(nflag2);
    if (lane == 0 && nflag2 != 0) atomicAdd(&pb.cnt->n_flag2, nflag2);
}

// ------------------------------------------------------------------------------------- setup


// Writes the powers tables of base x: tab[k] = first * x^k and tab[8 + k] = x^(8k), k < 8; returns
// x^64.  A lane's start first * x^tid then takes one product of two table words (two with x^64).
template <class F>
__device__ __forceinline__ typename F::T write_pow_tables(const F& f, typename F::T x, typename F::T first,
                                                         uint64_t* tab, unsigned long long& work) {
    typedef typename F::T T;
    T a = f.one();
    tab[0] = F::to_bits(first);
    for (int k = 0; k < 8; ++k) {
        if (k > 0) tab[k] = F::to_bits(f.mul(first, a));
        a = f.mul(a, x);
    }
    const T x8 = a;
    T c = f.one();
    for (int k = 0; k < 8; ++k) {
        tab[8 + k] = F::to_bits(c);
        c = f.mul(c, x8);
    }
    work += 23;
    return c;  // x^64
}


// x^L from x^64 for L = 64, 128 or 256 lanes.
template <class F>
__device__ __forceinline__ typename F::T lane_stride_pow(const F& f, typename F::T x64, uint32_t lanes) {
    typename F::T r = x64;
    for (uint32_t l = 64; l < lanes; l *= 2) r = f.mul(r, r);
    return r;
}


// Order mode: with g = gcd(Qo, p - 1) and e = (p - 1) / g, b^n == t forces (b^e)^n == t^e in the
// group of g-th roots of unity.  For each prime power r^a of g (Pohlig-Hellman), x = (b^e)^(g/r^a)
// and y = (t^e)^(g/r^a) give n == k_r (mod d_r), d_r = ord(x), or no solution when y is not a
// power of x (order_prime).  With n == cres (mod cm) the admissible n form one class
// n == n* (mod L), of which order_finish gives the first member n_start >= N0 and the count K below
// N1.  All moduli divide Qo < 2^32; only the exponent e needs 64-bit arithmetic.  The Fermat test to
// base b, b^(p-1) = (b^e)^g == 1, is part of the step for r = 2 (b mod p != 0: the filter hands
// p | b to the host).
// g = gcd(Qo, p - 1) from the valuations of p - 1 at the primes of Qo (prime powers in Qo are
// below 2^16); returns g and, 4 bits per prime, the exponent of the j-th prime in g.  Odd primes
// divide by multiplying with their inverse modulo 2^64 (no division instruction).
__device__ __forceinline__ uint64_t order_g(uint64_t p, const ItemParams& ip, uint64_t& gexp) {
    uint64_t g = 1;
    gexp = 0;
    for (uint32_t j = 0; j < ip.nopr; ++j) {
        const uint32_t r = ip.opr[j];
        uint32_t ra = 1, e = 0;
        if (r == 2) {
            const uint32_t v = __builtin_ctzll(p - 1);
            while (ra < ip.oprpow[j] && e < v) {
                ra *= 2;
                ++e;
            }
        } else {
            uint64_t x = p - 1;
            while (ra < ip.oprpow[j]) {
                const uint64_t y = x * ip.oinv[j];
                if (y > ip.olim[j]) break;  // r does not divide x
                x = y;
                ra *= r;
                ++e;
            }
        }
        gexp |= (uint64_t)e << (4 * j);
        g *= ra;
    }
    return g;
}




__device__ __forceinline__ uint64_t mulmod_qo(uint64_t a, uint64_t b, const ItemParams& ip) {
    return mulmod_q52(a, b, ip.Qo, ip.Qo_inv);
}

// Order mode, one round of 256 candidates (one per thread) as work lists: phase 1, a lane per
// candidate, computes g, b^e and t^e; phase 2, a lane per (candidate, prime of g) pair with the
// pairs binned by prime and exponent, runs that prime's Pohlig-Hellman step; order_finish, a lane
// per candidate again, adds the primes up by the CRT.  A wavefront of pairs shares one prime and
// one search bound, where a wavefront of candidates ran the step of every prime any of its lanes
// had, for as long as its costliest lane needed.
constexpr uint32_t kOrderRound = 256;
// Pair bins (bitlen(g) / 6, at most 3) * 96 + j * 6 + min(exponent, 5): a wavefront of pairs shares
// the prime and the search bound, and candidates with like g (so like power exponents h = g / r^a)
// come together, which the per-candidate class of g allows without a per-pair cost (a per-pair
// exact bit length of h measured slower: c2 setup 1.69 -> 1.83 ms).  On c2's pairs the lanes'
// loops run 44% of the wavefront's longest instead of 39%; c4 setup 19.7 -> 19.3 ms per item, and
// with two feeders c4 45.1 -> 42.4 ms (the setups of the two lanes then overlap each other less).
constexpr uint32_t kOrderBins = 4 * 6 * kOrderMaxPrimes;
static_assert(kOrderBins == 6 * 64, "wavefront 0 scans the pair bins, six per lane");
// Pair slots of a round: 8 per candidate, where a candidate averages about 2.5 (one pair per prime
// of Qo dividing p - 1); a round with more (never seen: over 70 standard deviations out) hands
// all its candidates to the host.
constexpr uint32_t kOrderPairs = 8 * kOrderRound;

template <class F>
struct OrderRound {
    uint64_t p[kOrderRound], aux[kOrderRound];     // the candidates' fields
    uint64_t one[F::kKeepsOne ? kOrderRound : 1];  // (the Montgomery field's 1)
    uint64_t Be[kOrderRound], Te[kOrderRound];     // b^e and t^e, e = (p - 1) / g
    uint64_t gexp[kOrderRound];                    // exponent of each prime of Qo in g, 4 bits each
    unsigned long long acc[kOrderRound];            // CRT sum over the primes
    unsigned long long lx[kOrderRound];             // exponent of each prime in L, 4 bits each
    uint64_t g[kOrderRound];                        // g; after phase 3 the candidate word
    uint8_t fail[kOrderRound];
    uint16_t pair[kOrderPairs];                     // candidate | prime index << 8
    uint32_t bin[kOrderBins];
    uint32_t npair, nkeep, base;
    uint32_t nxp, nq, na, nb, nkeep2;               // medium primes: pair slots, candidates, lists
};

// Medium primes (phases 3a-3c): pair slots of a round (kept index | table index << 8), in the pair
// array behind the list of kept candidates, with the results behind them; a candidate whose pairs
// do not fit takes none (dropping a congruence never drops a solution).  A round of 256 candidates
// has about 50 pairs at 1e12 (fewer than a wavefront, so they are not binned by prime).
constexpr uint32_t kExtPairs = 512;
static_assert(256 + 2 * kExtPairs + 256 <= kOrderPairs, "medium-prime pairs and list B fit behind the kept list");
constexpr uint16_t kExtFree = 0xFFFFu;    // slot of a candidate whose pairs did not fit
constexpr uint32_t kExtNone = 0xFFFFu;    // b^((p-1)/r) == 1 == t^((p-1)/r): no congruence
constexpr uint32_t kExtDead = 0xFFFEu;    // no n: t^((p-1)/r) is not a power of b^((p-1)/r)

// The medium primes 19 .. 1669 and their constants (r^-1 mod 2^64, floor((2^64 - 1) / r), 2^32 mod
// r, 1/r), read by uniform index as scalar loads.
__constant__ ExtTable c_ext = make_ext_table();

// v_readlane of a 64-bit value from lane j (j the same in every lane)
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)j);
    return (uint64_t)hi << 32 | lo;
}

// One lane's medium prime in phase 3a: the constants of the exact test "r divides x" (x > 0).  In
// the FP64 field (x < 2^50) x - rint(x * (1/r)) * r is zero exactly for multiples of r (rint gives
// x / r then, and the FMA is exact); the Montgomery field tests x * r^-1 mod 2^64 <=
// floor((2^64 - 1) / r).  A lane past the table's range holds r = 0 and 1/r = 0 (x - 0 is not zero),
// or r^-1 = 1 and a bound of 0 (x is not at most 0).
template <class F>
struct ExtTest {
    double a;         // FP64 field: 1/r
    uint32_t rr;      // and r
    uint64_t ia, ib;  // Montgomery field: r^-1 mod 2^64 and floor((2^64 - 1) / r)
    __device__ __forceinline__ void load(uint32_t e, uint32_t end) {
        const bool on = e < end;
        const ExtPrime& P = c_ext.e[on ? e : 0];
        if constexpr (F::kSmallMul) {
            a = on ? P.dinv : 0.0;
            rr = on ? P.r : 0u;
        } else {
            ia = on ? P.rinv : 1ull;
            ib = on ? P.lim : 0ull;
        }
    }
    __device__ __forceinline__ bool divides(uint64_t x, double xd) const {
        if constexpr (F::kSmallMul)
            return fma_rn(-rint(xd * a), (double)rr, xd) == 0.0;
        else
            return x * ia <= ib;
    }
};


// The sign of a small base's power: (-1)^e times the power of |s| when the base is negative.
__device__ __forceinline__ uint64_t mont_sign(const MontField& f, uint64_t x, uint64_t e, uint32_t neg) {
    return (neg && (e & 1)) ? mont_neg(f, x) : x;
}

// b^e and t^e (t = -c) modulo p in the field f: with the FP64 field, b and -c as small integers
// when their products with residues stay exact (fpow2_small); with the Montgomery field, |b| and
// |c| by windows when they are small (mont_pow2_small, ip.smw); else from their field forms.
template <class F>
__device__ __forceinline__ void pow_b_t(const F& f, uint64_t p, const ItemParams& ip, uint64_t e, typename F::T& Be,
                                        typename F::T& Te) {
    const uint64_t bmod = mod_signed(ip.b, p), tv = neg_mod(ip.c, p);
    if constexpr (F::kSmallMul) {
        const double sb = (double)bmod, sc = -(double)ip.c, pd = (double)p;
        if (sb * pd < 0x1p52 && fabs(sc) * pd < 0x1p52) {
            fpow2_small(f, sb, sc, e, Be, Te);
            return;
        }
    } else {
        if (ip.smw != 0) {
            mont_pow2_small(f, ip.bsp, ip.tsp, ip.smw, e, 1.0 / (double)p, Be, Te);
            Be = mont_sign(f, Be, e, ip.bneg);
            Te = mont_sign(f, Te, e, ip.tneg);
            return;
        }
    }
    const typename F::T rr = f.r2();
    fpow2(f, f.from(bmod, rr), f.from(tv, rr), e, Be, Te);
}

// log_x(y) for x of prime order r and y a power of x, by baby and giant steps in registers, shared
// by the S lanes of one pair (lane j of S, S a power of two): the low words of the keys of the B = 8
// baby values y * x^i (i < B), in every lane, against the giant values x^(B g), lane j taking
// g == j + 1 (mod S); about B + 3 + ceil(r / B) / S dependent products, where a plain search would
// chain r of them (more baby words spill registers in the order-mode setup).  A lane stops at its
// first word match (g, i) and, after the loop, every lane checks its match exactly (y * x^i against
// x^(B g), at most B - 1 more products, run once for the whole wavefront); the window of g holds
// just one log below r (the logs differ by r > B).  Returns the log when this lane found it, r when
// it found none (x is not of order r, which only a composite modulus allows) and r + 1 when its
// word match was not exact (the congruence is then not used; a miss of the right log is about
// 2^-32 per compared word).
template <class F>
__device__ __forceinline__ uint32_t small_log(const F& f, typename F::T x, typename F::T y, uint32_t r, uint32_t j,
                                              uint32_t S, unsigned long long& work) {
    typedef typename F::T T;
    constexpr uint32_t B = 8;
    const uint64_t ky = f.key(y);
    if (ky == f.key(f.one())) return j == 0 ? 0 : r;
    uint32_t bw[B];
    T v = y;
#pragma unroll
    for (uint32_t i = 0; i < B; ++i) {
        bw[i] = (uint32_t)f.key(v);
        if (i + 1 < B) v = f.mul(v, x);
    }
    T X = x;
    for (uint32_t q = 1; q < B; q <<= 1) X = f.mul(X, X);      // x^B
    T gv = X, XS = X;
    for (uint32_t q = 1; q < S; q <<= 1) XS = f.mul(XS, XS);  // x^(B S)
    for (uint32_t q = 0; q < j; ++q) gv = f.mul(gv, X);        // x^(B (j + 1))
    const uint32_t ng = (r + B - 1) / B;  // logs 1 .. r - 1 lie in B g - i, g <= ng, i < B
    work += B + 3;
    uint32_t g = j + 1, hit = B;
    for (; g <= ng; g += S) {
        const uint32_t w = (uint32_t)f.key(gv);
#pragma unroll
        for (uint32_t i = 0; i < B; ++i) hit = w == bw[i] ? i : hit;
        ++work;
        if (hit < B) break;
        gv = f.mul(gv, XS);
    }
    if (hit >= B) return r;
    T yx = y;  // y * x^hit, against gv = x^(B g)
    for (uint32_t i = 1; i < B; ++i)
        if (i <= hit) yx = f.mul(yx, x);
    const uint32_t k = g * B - hit;
    return f.key(yx) == f.key(gv) && k < r ? k : r + 1;
}

// t^e (tbase) or b^e modulo p, in the form pow_b_t picks for the pair (b, t): the small-integer
// form only when both bases allow it, so that the two lanes of a pair run the same instructions.
template <class F>
__device__ __forceinline__ typename F::T pow_b_or_t(const F& f, uint64_t p, const ItemParams& ip, uint64_t e, bool tbase) {
    const uint64_t bmod = mod_signed(ip.b, p), tv = neg_mod(ip.c, p);
    if constexpr (F::kSmallMul) {
        const double sb = (double)bmod, sc = -(double)ip.c, pd = (double)p;
        if (sb * pd < 0x1p52 && fabs(sc) * pd < 0x1p52) return fpow_small(f, tbase ? sc : sb, e);
    } else {
        if (ip.smw != 0) {
            uint32_t sp[4];
            for (uint32_t i = 0; i < 4; ++i) sp[i] = tbase ? ip.tsp[i] : ip.bsp[i];
            return mont_sign(f, mont_pow_small(f, sp, ip.smw, e, 1.0 / (double)p), e, tbase ? ip.tneg : ip.bneg);
        }
    }
    const typename F::T rr = f.r2();
    return fpow(f, f.from(tbase ? tv : bmod, rr), e);
}

// A value of lane `src` of the wavefront (every lane of the pair active).
template <class T>
__device__ __forceinline__ T lane_value(T v, uint32_t src) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)u, (int)src), hi = (uint32_t)__shfl((int)(uint32_t)(u >> 32), (int)src);
    return __builtin_bit_cast(T, (uint64_t)hi << 32 | lo);
}

// Phase 3b for the pair (candidate cc, medium prime r of p - 1), lane j of the pair's S lanes: with
// x = b^((p-1)/r) and y = t^((p-1)/r), b^n == t forces x^n == y.  x of order r (a prime) gives
// n == k = log_x(y) (mod r) (small_log), which with the candidate's class n == n_start (mod L) means
// n = n_start + L * J for J == (k - n_start) * L^-1 (mod r): the lane that finds k stores that
// residue of J.  x == 1 gives nothing when y == 1 and no n otherwise (stored by lane 0), as does an
// r that divides L already (a prime of cm) or an inexact word match.  The slot starts as kExtDead,
// so a search that fails (only a composite number can fail it) drops the number.
template <class F>
__device__ __forceinline__ void ext_log(const OrderRound<F>& R, uint32_t cc, const ExtPrime& P, const ItemParams& ip,
                                        uint32_t j, uint32_t S, uint16_t& res, unsigned long long& work) {
    typedef typename F::T T;
    F f;
    const uint64_t pc = R.p[cc];
    f.set(pc, R.aux[cc], F::kKeepsOne ? R.one[cc] : 0);
    const uint64_t h = (pc - 1) * P.rinv;  // (p - 1) / r, exact: r divides p - 1
    T x, y;
    if (S >= 2) {  // even lanes b^h, odd lanes t^h: one exponentiation's instructions for both
        const T v = pow_b_or_t(f, pc, ip, h, (j & 1u) != 0);
        const uint32_t lane = lane_id();
        x = lane_value(v, lane & ~1u);
        y = lane_value(v, lane | 1u);
    } else {
        pow_b_t(f, pc, ip, h, x, y);
    }
    if (j == 0) work += 2 * bitlen64(h);
    const uint64_t k1 = f.key(f.one());
    if (f.key(x) == k1) {
        if (j == 0) res = f.key(y) == k1 ? kExtNone : kExtDead;
        return;
    }
    const uint32_t k = small_log(f, x, y, P.r, j, S, work);
    if (k > P.r) res = kExtNone;
    if (k >= P.r) return;
    const uint32_t Lr = mod_small_f(R.acc[cc], P.r, P.r32, P.dinv);
    if (Lr == 0) {
        res = kExtNone;
        return;
    }
    const uint32_t nr = mod_small_f(R.lx[cc], P.r, P.r32, P.dinv);
    const uint32_t d = k >= nr ? k - nr : k + P.r - nr;
    res = (uint16_t)mod16_f(d * inv_mod_prime(Lr, P.r, P.dinv), P.r, P.dinv);
}

// Baby-step count m (whole rounds of the workgroup, at most mcap) and giant steps ng of an
// order-mode entry whose class has K exponents below N1 (jc = min(K, p): the exponents repeat with
// a period dividing p - 1).
__device__ __forceinline__ void order_sizes(uint64_t K, uint64_t p, const ItemParams& ip, uint64_t& jc, uint64_t& mg) {
    jc = K < p ? K : p;
    uint64_t m = isqrt_ceil64(jc);
    m = (m + ip.lanes - 1) / ip.lanes * ip.lanes;
    if (m > ip.mcap) m = ip.mcap;
    if (m == 0) m = 1;
    uint64_t rr;
    const uint64_t ng = (jc < (1ull << 51) ? divmod52(jc + m - 1, m, rr) : (jc + m - 1) / m) + 1;
    mg = m | (ng << 32);
}

// Phase 2 for the pair (candidate c, j-th prime r of Qo, r^a exactly dividing g): with h = g / r^a,
// x = (b^e)^h and y = (t^e)^h give n == k (mod d), d = ord(x), k = log_x(y); with n == cres modulo
// the r-part of cm the finer congruence holds, and its residue and the exponent of r in its modulus
// go to the candidate's CRT sum and exponent word.  For a prime x^(r^a) == b^(p-1) == 1, so d
// divides r^a; a number where that fails is composite and is dropped.  For r = 2, which divides
// every g, this is also the Fermat test to base b: x^(r^a) == (b^e)^g.
template <class F>
__device__ __forceinline__ void order_prime(OrderRound<F>& R, uint32_t c, uint32_t j, const ItemParams& ip,
                                            unsigned long long& work, uint32_t& composite) {
    typedef typename F::T T;
    F f;
    f.set(R.p[c], R.aux[c], F::kKeepsOne ? R.one[c] : 0);
    const uint32_t ej = (uint32_t)(R.gexp[c] >> (4 * j)) & 15u;
    const uint32_t r = ip.opr[j], rinv = (uint32_t)ip.oinv[j];
    // ra = r^ej and h = g / ra (exact), by multiplication only
    uint32_t ra = 1;
    uint64_t h = R.g[c];
    for (uint32_t k = 0; k < ej; ++k) {
        ra *= r;
        h = r == 2 ? h >> 1 : h * ip.oinv[j];
    }
    T x, y;
    fpow2(f, F::from_bits(R.Be[c]), F::from_bits(R.Te[c]), h, x, y);
    work += 2 * bitlen64(h);
    const uint64_t k1 = f.key(f.one()), ky = f.key(y);
    T xk = f.one();
    uint32_t kd = ~0u, dr = 0;
    for (uint32_t k = 0; k <= ra; ++k) {
        const uint64_t kx = f.key(xk);
        if (k > 0 && kx == k1) {
            dr = k;
            break;
        }
        if (kd == ~0u && kx == ky) kd = k;
        xk = f.mul(xk, x);
    }
    work += dr;
    // dr must be a power of r: divide ra by r down to dr, counting the exponent
    uint32_t q = ra, kq = ej;
    while (dr != 0 && q > dr) {
        q = r == 2 ? q >> 1 : q * rinv;
        --kq;
    }
    bool ok = dr != 0 && q == dr;
    if (!ok && j == 0) ++composite;  // (b^e)^g != 1: fails the Fermat test to base b
    ok = ok && kd != ~0u;             // else -c is not a power of b modulo p
    uint32_t cj = 0, kl = 0;
    if (ok) {
        const uint32_t cmj = ip.ocm[j], crj = ip.ocres[j];
        if (dr >= cmj) {
            ok = mod16(kd, cmj) == crj;
            cj = kd;
            kl = kq;
        } else {
            ok = mod16(crj, dr) == kd;
            cj = crj;
            for (uint32_t v = cmj; v > 1; v = r == 2 ? v >> 1 : v * rinv) ++kl;
        }
    }
    if (!ok) {
        R.fail[c] = 1u;  // every writer stores the same value
    } else {
        atomicAdd(&R.acc[c], (unsigned long long)mulmod_qo(cj, ip.ocrt[j], ip));  // below Qo
        atomicOr(&R.lx[c], (unsigned long long)kl << (4 * j));
    }
}

// The class of candidate c from its primes' results: n == n* (mod L); false when a prime failed or
// no n lies in [N0, N1).
template <class F>
__device__ __forceinline__ bool order_finish(const OrderRound<F>& R, uint32_t c, const ItemParams& ip, uint64_t& L,
                                             uint64_t& nst, uint64_t& K) {
    if (R.fail[c]) return false;
    uint64_t acc = R.acc[c];  // a sum of nopr terms below Qo: below 2^56
    const uint64_t gx = R.gexp[c], lx = R.lx[c];
    uint64_t Lw = 1;
    for (uint32_t j = 0; j < ip.nopr; ++j) {
        if ((gx >> (4 * j)) & 15u) {
            for (uint32_t k = (uint32_t)(lx >> (4 * j)) & 15u; k > 0; --k) Lw *= ip.opr[j];
        } else if (ip.ocm[j] > 1) {  // r does not divide g: n == cres modulo the r-part of cm alone
            acc += mulmod_qo(ip.ocres[j], ip.ocrt[j], ip);
            Lw *= ip.ocm[j];
        }
    }
    // acc modulo Qo (the quotient is below 2^5, and within one of the double one), then modulo L,
    // which divides Qo, from exact doubles
    int64_t aq = (int64_t)(acc - (uint64_t)((double)acc * ip.Qo_inv) * ip.Qo);
    while (aq < 0) aq += (int64_t)ip.Qo;
    while (aq >= (int64_t)ip.Qo) aq -= (int64_t)ip.Qo;
    uint64_t nstar, n0r;
    divmod52((uint64_t)aq, Lw, nstar);
    divmod52(ip.n0modQo, Lw, n0r);
    const uint64_t off = nstar >= n0r ? nstar - n0r : nstar + Lw - n0r;
    L = Lw;
    nst = ip.N0 + off;
    if (nst >= ip.N1) {
        K = 0;
        return false;
    }
    const uint64_t rest = ip.N1 - 1 - nst;
    uint64_t rr;
    K = (rest < (1ull << 52) ? divmod52(rest, Lw, rr) : rest / Lw) + 1;
    return true;
}

// The entry's step tables from B (baby base), G1 (giant start with one class, else 1), t and m:
// Bm = B^m, the start tables (baby starts t * B^k, giant starts G1 * Bm^k) and the lane strides.
template <class F>
__device__ __forceinline__ void write_step_tables(const F& f, uint64_t* E, typename F::T B, typename F::T G1,
                                                  typename F::T tm, uint64_t m, uint32_t lanes,
                                                  unsigned long long& work) {
    typedef typename F::T T;
    const T Bm = fpow(f, B, m);
    work += 2 * bitlen64(m);
    const T B64 = write_pow_tables(f, B, tm, E + E_BLO, work);
    const T Bm64 = write_pow_tables(f, Bm, G1, E + E_MLO, work);
    E[E_B64] = F::to_bits(B64);
    E[E_BM64] = F::to_bits(Bm64);
    E[E_BL] = F::to_bits(lane_stride_pow(f, B64, lanes));
    E[E_BML] = F::to_bits(lane_stride_pow(f, Bm64, lanes));
}

// Phase 4 of order mode: the entry of candidate cc of the round (its class n_start + k*L, k < jc,
// and its sizes from phase 3 or 3c) at entry slot `slot`, or a record for the host when the entry
// buffer is full.
template <class F>
__device__ __forceinline__ void order_entry(const OrderRound<F>& R, uint32_t cc, uint32_t slot, const ItemParams& ip,
                                            const PassBuffers& pb, unsigned long long& w_setup,
                                            unsigned long long& w_bsgs) {
    typedef typename F::T T;
    const uint64_t pc = R.p[cc];
    const uint32_t idx = (uint32_t)R.g[cc], info = (uint32_t)(R.g[cc] >> 32);
    if (slot >= pb.ent_cap) {  // no room: the host decides this one
        put_record(pb, atomicAdd(&pb.cnt->n_out, 1u), pb.out_index ? idx : pc, kHostDecides);
        return;
    }
    F f;
    f.set(pc, R.aux[cc], F::kKeepsOne ? R.one[cc] : 0);
    const uint64_t jc = R.Be[cc], mg = R.Te[cc], oL = R.acc[cc], onst = R.lx[cc];
    const uint64_t m = (uint32_t)mg;
    const T rr = f.r2();
    const T bm = f.from(mod_signed(ip.b, pc), rr), tm = f.from(neg_mod(ip.c, pc), rr);
    uint64_t* E = pb.ent + (uint64_t)slot * kEntWords;
    E[E_P] = pc;
    E[E_AUX] = f.aux();
    E[E_ONE] = F::to_bits(f.one());
    E[E_T] = F::to_bits(tm);
    E[E_MG] = mg;
    E[E_SSRC] = 1u | ((uint64_t)(info & 0xFFFFFFu) << 8) | ((uint64_t)idx << 32);
    T B, G1;
    bool small = false;
    if constexpr (!F::kSmallMul) {
        if (ip.smw != 0) {  // b^L and b^n_start by windows of the small base |b|
            const double pinv = 1.0 / (double)pc;
            B = mont_sign(f, mont_pow_small(f, ip.bsp, ip.smw, oL, pinv), oL, ip.bneg);
            G1 = mont_sign(f, mont_pow_small(f, ip.bsp, ip.smw, onst, pinv), onst, ip.bneg);
            small = true;
        }
    }
    if (!small) fpow_2e(f, bm, oL, onst, B, G1);
    w_setup += 2 * bitlen64(oL) + 2 * bitlen64(onst);
    E[E_G] = F::to_bits(G1);
    E[E_NB] = onst;
    E[E_META] = (1u << 11) | (1u << 12);
    E[E_KLIM] = jc;
    E[E_DJ] = oL;
    write_step_tables(f, E, B, G1, tm, m, ip.lanes, w_setup);
    w_bsgs += m + (mg >> 32);
}

template <class F, int MODE, bool EXT = false>
__device__ __forceinline__ void setup_pass(const PassDesc* __restrict__ d, PassBuffers& pb) {
    constexpr bool QMODE = MODE == kModeClass;
    const uint64_t lo = d->lo;
    const ItemParams& ip = d->ip;  // read in place: uniform fields load as scalars, tables by index
    typedef typename F::T T;
    __shared__ uint32_t sh[5];
    // class mode: each tile of kSetupTile candidates is processed in the order of the size of g, so
    // the lanes of a wavefront run loops of like length; order mode runs its rounds as work lists
    constexpr bool SORT = MODE == kModeClass;
    constexpr bool OMODE = MODE == kModeOrder;
    constexpr uint32_t kTile = SORT ? kSetupTile : 256;
    constexpr uint32_t kBins = 32;
    __shared__ uint16_t s_order[SORT ? kSetupTile : 1];
    __shared__ uint32_t s_bin[kBins];
    typedef typename std::conditional<OMODE, OrderRound<F>, uint32_t>::type RoundStore;
    __shared__ RoundStore s_round;
    const uint32_t n = min(pb.cnt->n_cand, pb.cand_cap);
    const uint32_t lane = lane_id();
    unsigned long long w_fermat = 0, w_setup = 0, w_bsgs = 0;
    uint32_t fermat_fail = 0;
    for (uint32_t t0 = blockIdx.x * kTile; t0 < n; t0 += gridDim.x * kTile) {
        if (SORT) {
            if (threadIdx.x < kBins) s_bin[threadIdx.x] = 0;
            __syncthreads();
            uint32_t keys[kTile / 256];
            for (uint32_t q = 0; q < kTile / 256; ++q) {
                const uint32_t i = t0 + q * 256 + threadIdx.x;
                uint32_t key = kBins - 1;  // past the end: last
                if (i < n) {
                    const uint32_t ix = (uint32_t)pb.cand[i];
                    const uint64_t pk = pb.vals ? pb.vals[ix] : lo + 2 * (uint64_t)ix;
                    key = min(exp_classes(pk, ip).g >> 1, 30u);
                }
                keys[q] = key;
                atomicAdd(&s_bin[key], 1u);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t tot = 0;
                for (uint32_t b = 0; b < kBins; ++b) {
                    const uint32_t c = s_bin[b];
                    s_bin[b] = tot;
                    tot += c;
                }
            }
            __syncthreads();
            for (uint32_t q = 0; q < kTile / 256; ++q) s_order[atomicAdd(&s_bin[keys[q]], 1u)] = (uint16_t)(q * 256 + threadIdx.x);
            __syncthreads();
        }
        if constexpr (OMODE) {
            OrderRound<F>& R = s_round;
            const uint32_t c = threadIdx.x, i = t0 + c;
            const bool live = i < n;
            uint64_t gx = 0;
            uint32_t gcls = 0;  // bitlen(g) / 6, at most 3: the pairs' powers have exponents g / r^a
            for (uint32_t b = c; b < kOrderBins; b += 256) R.bin[b] = 0;  // more bins than threads
            if (c == 0) R.nkeep = 0;
            // phase 1: g, b^e and t^e of each candidate
            if (live) {
                const uint32_t ix = (uint32_t)pb.cand[i];
                const uint64_t pc = pb.vals ? pb.vals[ix] : lo + 2 * (uint64_t)ix;
                F f;
                f.init(pc);
                const uint64_t g = order_g(pc, ip, gx);
                gcls = min(bitlen64(g) / 6u, 3u) * 96u;
                const uint32_t tz = __builtin_ctzll(g);
                const uint64_t e = exact_div_odd((pc - 1) >> tz, g >> tz);
                T Be, Te;
                pow_b_t(f, pc, ip, e, Be, Te);
                w_setup += 2 * bitlen64(e);
                R.p[c] = pc;
                R.aux[c] = f.aux();
                if (F::kKeepsOne) R.one[c] = F::to_bits(f.one());
                R.Be[c] = F::to_bits(Be);
                R.Te[c] = F::to_bits(Te);
                R.g[c] = g;
                R.gexp[c] = gx;
            }
            R.fail[c] = live ? 0 : 1;
            R.acc[c] = 0;
            R.lx[c] = 0;
            __syncthreads();
            // the (candidate, prime) pairs, binned by prime and exponent (a counting sort)
            if (live)
                for (uint32_t j = 0; j < ip.nopr; ++j) {
                    const uint32_t ej = (uint32_t)(gx >> (4 * j)) & 15u;
                    if (ej != 0) atomicAdd(&R.bin[gcls + j * 6 + min(ej, 5u)], 1u);
                }
            __syncthreads();
            if (c < 64) {  // wavefront 0: exclusive prefix sum of the bins, six per lane
                uint32_t v[6], s = 0;
#pragma unroll
                for (uint32_t k = 0; k < 6; ++k) {
                    v[k] = R.bin[6 * c + k];
                    s += v[k];
                }
                uint32_t incl = s;
#pragma unroll
                for (uint32_t o = 1; o < 64; o <<= 1) {
                    const uint32_t u = (uint32_t)__shfl_up((int)incl, o);
                    if (c >= o) incl += u;
                }
                uint32_t ex = incl - s;
#pragma unroll
                for (uint32_t k = 0; k < 6; ++k) {
                    R.bin[6 * c + k] = ex;
                    ex += v[k];
                }
                if (c == 63) R.npair = incl;
            }
            __syncthreads();
            const bool overflow = R.npair > kOrderPairs;  // the same in every lane
            if (live && !overflow)
                for (uint32_t j = 0; j < ip.nopr; ++j) {
                    const uint32_t ej = (uint32_t)(gx >> (4 * j)) & 15u;
                    if (ej != 0) R.pair[atomicAdd(&R.bin[gcls + j * 6 + min(ej, 5u)], 1u)] = (uint16_t)(c | (j << 8));
                }
            if (live && overflow) {  // no room for the round's pairs: the host decides its candidates
                put_record(pb, atomicAdd(&pb.cnt->n_out, 1u), pb.out_index ? (uint32_t)pb.cand[i] : R.p[c], kHostDecides);
                R.fail[c] = 1u;
            }
            __syncthreads();
            // phase 2: a lane per pair
            const uint32_t np = overflow ? 0u : R.npair;
            for (uint32_t k = c; k < np; k += 256) {
                const uint32_t pr = R.pair[k];
                order_prime(R, pr & 255u, pr >> 8, ip, w_setup, fermat_fail);
            }
            __syncthreads();
            // phase 3: the class of each candidate; the kept ones are listed (in the pair array, free
            // now) with their sizes in the round's arrays
            {
                uint64_t oL = 0, onst = 0, K = 0;
                if (live && order_finish(R, c, ip, oL, onst, K)) {
                    uint64_t jc, mg;
                    order_sizes(K, R.p[c], ip, jc, mg);
                    R.Be[c] = jc;
                    R.Te[c] = mg;
                    R.acc[c] = oL;
                    R.lx[c] = onst;
                    R.g[c] = pb.cand[i];  // index | info << 32
                    R.pair[atomicAdd(&R.nkeep, 1u)] = (uint16_t)c;
                }
            }
            __syncthreads();
            const uint32_t nk = R.nkeep;
            // medium primes of p - 1 (phases 3a-3c, the kernel variant EXT) narrow the classes of at
            // least ext_kmin exponents; the entries are then the candidates that phase 3c keeps, listed
            // in R.fail
            const bool ext = EXT && nk != 0;  // the same in every lane
            if constexpr (EXT) if (ext) {
                uint16_t* xp = R.pair + 256;     // pair slots: kept index | table index << 8
                uint16_t* xr = xp + kExtPairs;   // per slot: n mod r, kExtNone or kExtDead
                if (c == 0) {
                    R.nxp = 0;
                    R.nq = 0;
                    R.na = 0;
                    R.nb = 0;
                    R.nkeep2 = 0;
                }
                __syncthreads();
                // the kept candidates whose class has at least ext_kmin exponents, listed (in R.fail,
                // free until 3c) so that the trial divisions run on whole wavefronts of them
                for (uint32_t k = c; k < nk; k += 256) {
                    R.gexp[k] = 0;  // no pairs
                    if (R.Be[R.pair[k]] >= ip.ext_kmin) R.fail[atomicAdd(&R.nq, 1u)] = (uint8_t)k;
                }
                __syncthreads();
                // 3a, a wavefront per listed candidate and a lane per medium prime: lane l tests the table's
                // primes ext_first + l + 64 u (u < 3), whose constants it loads once a round; the
                // wavefront's candidates (listed ones wv, wv + 4, ...) are gathered one per lane and
                // taken one by one with v_readlane, and the primes that divide p - 1 come back as a
                // ballot.  A candidate keeps at most 8 of them, ascending, while L times their product
                // stays below 2^62 (so that n_start + L * j and L * r fit).
                {
                    const uint32_t nq = R.nq, wv = c >> 6, ln = c & 63u;
                    ExtTest<F> tests[3];
#pragma unroll
                    for (uint32_t u = 0; u < 3; ++u) tests[u].load(ip.ext_first + ln + 64 * u, ip.ext_end);
                    const uint32_t qj = wv + 4 * ln;  // this lane's candidate of the wavefront
                    uint64_t xj = 0, Lj = 0;
                    uint32_t kj = 0;
                    if (qj < nq) {
                        kj = R.fail[qj];
                        const uint32_t cj = R.pair[kj];
                        xj = R.p[cj] - 1;
                        Lj = R.acc[cj];
                    }
                    const uint32_t nw = nq > wv ? (nq - wv + 3) / 4 : 0;  // the same in the wavefront
                    for (uint32_t j = 0; j < nw; ++j) {
                        const uint64_t x = readlane64(xj, j);
                        uint64_t lp = readlane64(Lj, j);  // L times the primes taken
                        const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)kj, (int)j);
                        const double xd = (double)x;       // exact in the FP64 field (p < 2^50)
                        uint64_t es = 0;
                        uint32_t ne = 0;
#pragma unroll
                        for (uint32_t u = 0; u < 3; ++u) {
                            uint64_t m = __ballot(tests[u].divides(x, xd));
                            while (m != 0) {  // ascending primes
                                const uint32_t e = ip.ext_first + 64 * u + (uint32_t)__builtin_ctzll(m);
                                m &= m - 1;
                                const ExtPrime P = c_ext.e[e];
                                if (lp <= (P.lim >> 2) && ne < 8) {
                                    es |= (uint64_t)e << (8 * ne);
                                    ++ne;
                                    lp *= P.r;
                                }
                            }
                        }
                        uint32_t start = 0;
                        if (ne != 0) {
                            if (ln == 0) start = atomicAdd(&R.nxp, ne);
                            start = (uint32_t)__builtin_amdgcn_readfirstlane((int)start);
                            const bool fits = start + ne <= kExtPairs;
                            if (ln < ne && start + ln < kExtPairs) {
                                const uint32_t e = (uint32_t)(es >> (8 * ln)) & 255u;
                                xp[start + ln] = fits ? (uint16_t)(k | e << 8) : kExtFree;
                                xr[start + ln] = (uint16_t)kExtDead;
                            }
                            if (!fits) ne = 0;
                        }
                        if (ln == 0) R.gexp[k] = start | (uint64_t)ne << 32;
                    }
                }
                __syncthreads();
                // the kept candidates without pairs are final (list A, in R.fail: their entries are
                // written now, beside the pairs' searches); the others (list B, kept indices behind the
                // pair slots) wait for 3c
                uint16_t* lb = xr + kExtPairs;
                for (uint32_t k = c; k < nk; k += 256) {
                    if ((R.gexp[k] >> 32) == 0)
                        R.fail[atomicAdd(&R.na, 1u)] = (uint8_t)R.pair[k];
                    else
                        lb[atomicAdd(&R.nb, 1u)] = (uint16_t)k;
                }
                __syncthreads();
                const uint32_t na = R.na;
                if (c == 0) R.base = na ? atomicAdd(&pb.cnt->n_ent, na) : 0u;
                __syncthreads();
                // 3b, S lanes per pair (the lanes of a pair share one wavefront and each computes the
                // pair's powers): in the FP64 field, whose searches are latency-bound, up to 4 while
                // they fit the workgroup (a round has about 30 pairs at 1e12); in the Montgomery field,
                // whose 62-bit powers make the pair instruction-bound, one (measured: c2 setup 2.12
                // ms per item with one lane, 2.01 with four; c4 22.7 ms with one, 23.9 with four).
                // The pairs take the first wavefronts and list A's entries the others, so that the
                // two kinds of dependent chains run side by side.
                {
                    const uint32_t nxs = min(R.nxp, kExtPairs);
                    const uint32_t lsh = F::kSmallMul ? (nxs <= 64 ? 2 : nxs <= 128 ? 1 : 0) : (nxs <= 128 ? 1 : 0);
                    const uint32_t pt = nxs << lsh, wp = min(4u, (pt + 63) / 64), ew = 4 - wp;
                    if (c < wp * 64) {
                        for (uint32_t t = c; t < pt; t += wp * 64) {
                            const uint32_t s = t >> lsh, v = xp[s];
                            if (v != kExtFree)
                                ext_log(R, R.pair[v & 255u], c_ext.e[v >> 8], ip, t & ((1u << lsh) - 1), 1u << lsh, xr[s],
                                        w_setup);
                        }
                    }
                    if (ew == 0 || c >= wp * 64) {
                        const uint32_t base = R.base;
                        for (uint32_t q = ew ? c - wp * 64 : c; q < na; q += ew ? ew * 64 : 256)
                            order_entry(R, R.fail[q], base + q, ip, pb, w_setup, w_bsgs);
                    }
                }
                __syncthreads();
                // 3c, a lane per candidate of list B: its primes' residues J == J_r (mod r) combine
                // into J == J0 (mod M) (narrow_class on the pair (M, J0); M is the product of the
                // primes, J0 < M), and the class n == n_start (mod L) narrows to n == n_start + L * J0
                // (mod L * M), whose first member from N0 is n_start + L * J0
                const uint32_t nb = R.nb;
                for (uint32_t qb = c; qb < nb; qb += 256) {
                    const uint32_t k = lb[qb], cc = R.pair[k];
                    const uint64_t se = R.gexp[k];
                    const uint32_t start = (uint32_t)se, ne = (uint32_t)(se >> 32);
                    bool alive = true;
                    uint64_t M = 1, J0 = 0;
                    for (uint32_t q = 0; q < ne; ++q) {
                        const uint32_t s = start + q, res = xr[s];
                        if (res == kExtDead) {
                            alive = false;
                            break;
                        }
                        if (res == kExtNone) continue;
                        const ExtPrime P = c_ext.e[xp[s] >> 8];
                        if (M == 1) {
                            M = P.r;
                            J0 = res;
                        } else {
                            narrow_class(M, J0, res, P.r, P.r32, P.dinv);
                        }
                    }
                    uint64_t L = R.acc[cc], nst = R.lx[cc];
                    const bool narrowed = M > 1;
                    nst += L * J0;
                    L *= M;
                    if (alive && narrowed) {
                        if (nst >= ip.N1) {
                            alive = false;
                        } else {
                            const uint64_t rest = ip.N1 - 1 - nst;
                            uint64_t rr;
                            const uint64_t K = (L > rest ? 0 : rest < (1ull << 52) ? divmod52(rest, L, rr) : rest / L) + 1;
                            uint64_t jc, mg;
                            order_sizes(K, R.p[cc], ip, jc, mg);
                            R.Be[cc] = jc;
                            R.Te[cc] = mg;
                            R.acc[cc] = L;
                            R.lx[cc] = nst;
                        }
                    }
                    if (alive) R.fail[atomicAdd(&R.nkeep2, 1u)] = (uint8_t)cc;
                }
                __syncthreads();
            }
            const uint32_t nent = ext ? R.nkeep2 : nk;
            if (c == 0) R.base = nent ? atomicAdd(&pb.cnt->n_ent, nent) : 0u;  // one reservation a round
            __syncthreads();
            // phase 4: the entries, a lane per kept candidate (a quarter of the candidates or so: a
            // lane per candidate left most lanes idle through the entry's hundred-odd products)
            const uint32_t base = R.base;
            for (uint32_t k = c; k < nent; k += 256)
                order_entry(R, ext ? R.fail[k] : R.pair[k], base + k, ip, pb, w_setup, w_bsgs);
        }
        for (uint32_t q = 0; q < (OMODE ? 0 : kTile / 256); ++q) {
            const uint32_t li = SORT ? (uint32_t)s_order[q * 256 + threadIdx.x] : q * 256 + threadIdx.x;
            const uint32_t i = t0 + li;
            bool keep = false;
            uint64_t p = 0, m = 0, ng = 0, jc = 0;
            uint32_t idx = 0, info = 0, s = 1, mask = 0;
            ExpClasses ec{0, 0, 0};
            F f;
            if (i < n) {
                const uint64_t cd = pb.cand[i];
                idx = (uint32_t)cd;
                info = (uint32_t)(cd >> 32);
                p = pb.vals ? pb.vals[idx] : lo + 2 * (uint64_t)idx;
                f.init(p);
                // order mode runs its Fermat test inside order_prime, on the power b^e it needs anyway
                const bool fermat_here = pb.need_prime && MODE != kModeOrder;
                if (fermat_here) w_fermat += bitlen64(p - 1);
                if (fermat_here && !fermat2(f, p)) {
                    ++fermat_fail;
                } else {
                    const uint64_t bmod = mod_signed(ip.b, p), t = neg_mod(ip.c, p);
                    const T rr = f.r2();
                    const T bm = f.from(bmod, rr), tm = f.from(t, rr);
                    if constexpr (QMODE) {
                        ec = exp_classes(p, ip);
                        mask = info;
                        if (ec.g > 2) {
                            // b^n == t forces w^n == u for w = b^e, u = t^e, e = (p - 1) / g, in the
                            // group of g-th roots of unity: n == k0 (mod d) for d = ord(w) (a divisor
                            // of g) and k0 = log_w(u), or no n at all.  One scan of w^k finds both.
                            const uint64_t e = (p - 1) / ec.g;
                            T u, w;
                            fpow2(f, tm, bm, e, u, w);
                            const uint64_t ku = f.key(u), k1 = f.key(f.one());
                            uint32_t d = 0, k0 = ~0u;
                            T wk = f.one();
                            for (uint32_t k = 0; k <= ec.g; ++k) {
                                const uint64_t kw = f.key(wk);
                                if (k > 0 && kw == k1) {
                                    d = k;
                                    break;
                                }
                                if (k0 == ~0u && kw == ku) k0 = k;
                                wk = f.mul(wk, w);
                            }
                            w_setup += 2 * bitlen64(e) + d;
                            // class k keeps n == n_start + k*cm (mod qp): n mod d from n_start mod Q
                            uint32_t keepm = 0;
                            if (d != 0 && k0 != ~0u) {
                                const uint32_t step = (uint32_t)ip.cm % d;  // cm divides Q <= 48
                                uint32_t r = ip.nsmodQ % d;
                                for (uint32_t k = 0; k < ec.ncls; ++k) {
                                    if (r == k0) keepm |= 1u << k;
                                    r += step;
                                    r = r >= d ? r - d : r;
                                }
                            }
                            mask &= keepm;  // d == 0: b^(p-1) != 1, a composite
                        }
                        if (mask != 0) {
                            s = __popc(mask);
                            const uint2 sz = ip.csz[ec.qp * 32 + s];
                            m = sz.x;
                            ng = sz.y;
                            keep = true;
                        }
                    } else {
                        const uint64_t g = gcd64(ip.cm, pm1_mod(p, ip.cm, ip.cm_r32, ip.cm_dinv));
                        bool ok = true;
                        if (g > 2) {
                            // B = b^cm is a g-th power, so T = t*b^-n_start must be one: t^e == G0^e
                            const uint64_t e = (p - 1) / g;
                            const T G0 = fpow(f, bm, ip.n_start);
                            T te, ge;
                            fpow2(f, tm, G0, e, te, ge);
                            ok = f.key(te) == f.key(ge);
                            w_setup += 2 * bitlen64(e) + bitlen64(ip.n_start);
                        }
                        const uint32_t stride = (info & 1u) ? 2 : 1, parity = (info >> 1) & 1u;
                        const uint64_t kc = ip.kcount < p ? ip.kcount : p;
                        if (ok && kc > parity) {
                            const uint32_t combo = (stride - 1) * 2 + parity;
                            if (kc == ip.kcount) {
                                jc = ip.pjc[combo];
                                m = ip.pm[combo];
                                ng = ip.png[combo];
                            } else {  // kcount capped at p (p below the exponent count)
                                jc = (kc - parity + stride - 1) / stride;
                                m = isqrt_ceil64(jc);
                                if (m > ip.mcap) m = ip.mcap;
                                if (m == 0) m = 1;
                                ng = (jc + m - 1) / m + 1;  // g = 0 .. ceil(jcount/m)
                            }
                            keep = true;
                        }
                    }
                }
            }
            const uint32_t slot = block_append<4>(keep, &pb.cnt->n_ent, sh);
            if (keep && slot >= pb.ent_cap) {
                const uint32_t o = atomicAdd(&pb.cnt->n_out, 1u);  // no room: the host decides this one
                put_record(pb, o, pb.out_index ? idx : p, kHostDecides);
                keep = false;
            }
            if (keep) {
                uint64_t* E = pb.ent + (uint64_t)slot * kEntWords;
                const uint64_t bmod = mod_signed(ip.b, p), t = neg_mod(ip.c, p);
                const T rr = f.r2();
                const T bm = f.from(bmod, rr), tm = f.from(t, rr);
                T B;
                T G1 = f.one();  // the giant start when the entry has one class
                E[E_P] = p;
                E[E_AUX] = f.aux();
                E[E_ONE] = F::to_bits(f.one());
                E[E_T] = F::to_bits(tm);
                E[E_MG] = m | (ng << 32);
                E[E_SSRC] = (uint64_t)s | ((uint64_t)((QMODE ? mask : info) & 0xFFFFFFu) << 8) | ((uint64_t)idx << 32);
                if (QMODE) {
                    // baby base B = b^qp; class k starts at n_start + k*cm, so its giant start is
                    // G0 * C^k with G0 = b^n_start and C = b^cm (B and G0 share their squarings)
                    T G0;
                    fpow_2e(f, bm, ec.qp, ip.n_start, B, G0);
                    const T C = fpow(f, bm, ip.cm);
                    E[E_NB] = ip.N0;
                    E[E_META] = ec.qp | (1u << 10);
                    E[E_KLIM] = 0;
                    w_setup += 2 * bitlen64(ip.n_start) + 2 * bitlen64(ip.cm) + 12;
                    // per kept class, in ascending k: the giant start and (8 bits each) the offset of
                    // its first exponent from N0, n_start - N0 + k*cm < qp
                    const uint32_t off0 = (uint32_t)(ip.n_start - ip.N0), cm = (uint32_t)ip.cm;
                    uint64_t offs[3] = {0, 0, 0};
                    uint32_t o = 0;
                    T Gk = G0;
                    for (uint32_t k = 0; (mask >> k) != 0; ++k) {
                        if ((mask >> k) & 1u) {
                            E[E_G + o] = F::to_bits(Gk);
                            offs[o >> 3] |= (uint64_t)(off0 + k * cm) << (8 * (o & 7));
                            ++o;
                        }
                        Gk = f.mul(Gk, C);
                        ++w_setup;
                    }
                    E[E_DJ] = offs[0];
                    E[E_DJ + 1] = offs[1];
                    E[E_DJ + 2] = offs[2];
                } else {
                    T G0;
                    fpow_2e(f, bm, ip.cm, ip.n_start, B, G0);
                    w_setup += 2 * bitlen64(ip.n_start) + 2 * bitlen64(ip.cm);
                    if (info & 1u) {  // stride 2: k = 2j + parity
                        if (info & 2u) G0 = f.mul(G0, B);
                        B = f.mul(B, B);
                    }
                    G1 = G0;
                    E[E_G] = F::to_bits(G0);
                    E[E_NB] = ip.n_start;
                    E[E_META] = ((info & 1u) ? (1u << 8) : 0u) | ((info & 2u) ? (1u << 9) : 0u);
                    E[E_KLIM] = jc;
                }
                // baby starts t * B^tid; with one class (plain mode) the giant starts G * (B^m)^tid
                // as well (E_META bit 12)
                write_step_tables(f, E, B, G1, tm, m, ip.lanes, w_setup);
                if (!QMODE) E[E_META] |= 1u << 12;
                w_bsgs += m + (uint64_t)s * ng;
            }
        }
        if (SORT || OMODE) __syncthreads();  // the tile's LDS is rewritten by the next tile
    }
    // one atomic per counter and workgroup
    __shared__ unsigned long long red[4][4];
    const uint32_t w = threadIdx.x >> 6;
    w_fermat = wave_sum64(w_fermat);
    w_setup = wave_sum64(w_setup);
    w_bsgs = wave_sum64(w_bsgs);
    const unsigned long long ff = wave_sum64(fermat_fail);
    if (lane == 0) {
        red[0][w] = w_fermat;
        red[1][w] = w_setup;
        red[2][w] = w_bsgs;
        red[3][w] = ff;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const unsigned long long v = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
        if (v != 0) {
            if (threadIdx.x == 0) atomicAdd(&pb.cnt->w_fermat, v);
            if (threadIdx.x == 1) atomicAdd(&pb.cnt->w_setup, v);
            if (threadIdx.x == 2) atomicAdd(&pb.cnt->w_bsgs, v);
            if (threadIdx.x == 3) atomicAdd(&pb.cnt->n_fermat_fail, (uint32_t)v);
        }
    }
}

template <class F, int MODE>
__global__ __launch_bounds__(256) void k_setup(const PassDesc* __restrict__ d, PassBuffers pb) {
    setup_pass<F, MODE>(d, pb);
}

// Order mode at F::kOrderSetupWaves waves per SIMD (FP64 field: 7, i.e. 72 registers, measured
// faster than 82 registers at 5 waves, c3 setup 1.23 -> 1.13 ms; the Montgomery field's LDS allows
// 6); the other modes keep their registers.
template <class F, bool EXT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(F::kOrderSetupWaves))) void k_setup_order(
    const PassDesc* __restrict__ d, PassBuffers pb) {
    setup_pass<F, kModeOrder, EXT>(d, pb);
}

// ------------------------------------------------------------------------------------- bsgs
//
// Table (LDS, one per entry): 2^nbbits buckets of 8 slots.  A slot is a 16-bit fingerprint (one
// 16-byte bucket = one 128-bit LDS read); one u32 fill counter per bucket; no baby indices.  A baby
// step takes the next slot of its home bucket from the counter (slots fill in order, so a bucket is
// full exactly when slot 7 is written) and spills to the next bucket when full.  A lookup reads its
// home bucket and continues only while buckets are full.  Every fingerprint match is verified
// exactly after the giant loop (the baby values t * B^i are run again and compared with the giant
// value), so the device verdicts are exact.
//
// Hash of a residue from its field's 32-bit hash word w (F::hash_word): fingerprint = w & 0xFFFF
// (0xFFFF marks an empty slot, so it maps to 0xFFFE), bucket = (w >> 20) & (2^nbbits - 1), carried as
// the bucket's byte offset (w >> 16) & (mask << 4).  Fingerprint and bucket come from disjoint bits
// of w.

constexpr uint32_t kEmpty16 = 0xFFFFu;

struct BsgsTable {
    uint16_t* fp;    // [nbk * 8]
    uint32_t* cnt;   // [nbk]
    uint32_t omask;  // (nbk - 1) << 4: bucket byte offsets
};

__device__ __forceinline__ void hash_word(uint32_t w, uint32_t omask, uint32_t& fp, uint32_t& boff) {
    fp = min(w & 0xFFFFu, 0xFFFEu);
    boff = (w >> 16) & omask;
}

// boff is a multiple of 16 (a bucket's byte offset), so the bucket's counter is at byte offset
// boff / 4 of the counters and slot k at byte offset boff + 2k of the fingerprints; byte offsets
// keep the compiler from masking the low bits again (two address operations instead of five).
__device__ __forceinline__ void tab_insert(const BsgsTable& T, uint32_t boff, uint32_t fp) {
    for (;;) {
        uint32_t* c = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(T.cnt) + (boff >> 2));
        const uint32_t old = atomicAdd(c, 1u);
        if (old < 8) {
            *reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(T.fp) + boff + 2u * old) = (uint16_t)fp;
            return;
        }
        boff = (boff + 16u) & T.omask;
    }
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const u32x4 lds_u32x4;
typedef __attribute__((address_space(3))) uint16_t lds_u16;

// The bucket at byte offset boff.  The table starts the kernel's dynamic LDS, which starts at LDS
// address 0 because k_bsgs has no static LDS (bsgs_lds_at_zero checks that on the host), so boff is
// the bucket's LDS address: the giant loop needs no add for the table base.
__device__ __forceinline__ uint4 read_bucket(const BsgsTable&, uint32_t boff) {
    const u32x4 w = *reinterpret_cast<lds_u32x4*>((size_t)boff);
    return make_uint4(w.x, w.y, w.z, w.w);
}


template <class F>
struct BsgsEntry {
    typedef typename F::T T;
    F f;
    const uint64_t* E;
    T t, B64;
    uint32_t m, ng, s, meta, qp;
    uint64_t nb, klim;
};

// Giant steps whose fingerprint is in the table, verified after the giant loop (kept out of the
// loop so that the exact check's registers do not count against the loop's occupancy).  A hit
// holds the giant value's key, the giant index g and the class j; the table keeps no baby
// indices, so the check runs the baby steps again and compares every baby value with every hit's
// key exactly.  More than kMaxHits hits in one entry (a short cycle of the giant values, say)
// leave the number to the host.
constexpr uint32_t kMaxHits = 16;
constexpr uint32_t kCycleFlag = 1u << 30;  // hit count of an entry whose baby values repeat

struct HitList {
    uint4 h[kMaxHits];
    uint32_t n;
    uint32_t ovf;  // a baby step found its bucket full and went on along the chain
    unsigned long long best;
};

// A giant value whose fingerprint is somewhere in the chain from bucket goes to the hit list once.
__device__ __forceinline__ void collect_hit(const BsgsTable& T, uint32_t boff, uint32_t fp, uint64_t xv, uint32_t g,
                                            uint32_t j, HitList& hl) {
    for (uint32_t b = boff;; b = (b + 16u) & T.omask) {
        const uint4 w = read_bucket(T, b);
        if (match_mask(w, fp) != 0) {
            const uint32_t slot = atomicAdd(&hl.n, 1u);
            if (slot < kMaxHits) hl.h[slot] = make_uint4((uint32_t)xv, (uint32_t)(xv >> 32), g, j);
            return;
        }
        if (w.w >= 0xFFFF0000u) return;  // slot 7 empty: the chain ends here
    }
}

// x^tid for tid < 64 * WB from the entry's tables: x^(tid & 63) * (x^64)^(tid >> 6); the lane
// stride x^(64 * WB) comes precomputed from the entry (word stride_word).
template <class F, int WB>
__device__ __forceinline__ typename F::T thread_pow(const F& f, const uint64_t* E, uint32_t lo_word, uint32_t hi_word,
                                                  typename F::T x64, uint32_t stride_word, uint32_t tid,
                                                  typename F::T& stride_pow) {
    typedef typename F::T T;
    T r = f.mul(F::from_bits(E[lo_word + (tid & 7)]), F::from_bits(E[hi_word + ((tid >> 3) & 7)]));
    const uint32_t w = tid >> 6;
    if (WB >= 2 && (w & 1)) r = f.mul(r, x64);
    if (WB >= 4 && (w & 2)) r = f.mul(r, f.mul(x64, x64));
    stride_pow = F::from_bits(E[stride_word]);
    return r;
}

// One workgroup of WB wavefronts per entry: the table lives in LDS once per entry and every lane
// of the group takes baby steps i == tid and giant steps g == tid (mod 64 * WB).
template <class F, int WB>
__global__ __launch_bounds__(64 * WB) void k_bsgs(const PassDesc* __restrict__ d, PassBuffers pb) {
    const uint64_t item_N1 = d->ip.N1, item_cm = d->ip.cm;
    const TableGeom tg = d->tg;
    typedef typename F::T T;
    constexpr uint32_t L = 64 * WB;
    // dynamic LDS: the table (20 bytes per bucket) at offset 0, so that a bucket's byte offset is its
    // LDS address, then two hit lists used by alternate entries
    extern __shared__ uint32_t lds_tab[];
    const uint32_t tid = threadIdx.x;
    const uint32_t nbk = 1u << tg.nbbits;
    HitList* hitl = reinterpret_cast<HitList*>(lds_tab + 5u * nbk);
    BsgsTable Tb;
    Tb.fp = reinterpret_cast<uint16_t*>(lds_tab);
    Tb.cnt = reinterpret_cast<uint32_t*>(Tb.fp + nbk * 8u);
    uint4* fp4 = reinterpret_cast<uint4*>(Tb.fp);
    const uint32_t n = min(pb.cnt->n_ent, pb.ent_cap);
    // the table starts empty; afterwards every entry empties the buckets it used before the barrier
    // that ends it, so the next entry needs no barrier of its own between the clear and its inserts
    for (uint32_t i = tid; i < nbk; i += L) {
        fp4[i] = make_uint4(~0u, ~0u, ~0u, ~0u);
        Tb.cnt[i] = 0u;
    }
    if (tid == 0) {
        hitl[0].n = 0;
        hitl[0].ovf = 0;
        hitl[0].best = ~0ull;
    }
    __syncthreads();
    uint32_t par = 0;
    // entries blockIdx.x, blockIdx.x + gridDim.x, ...: no shared claim counter (one same-address
    // atomic per entry costs more than a short entry's whole BSGS)
    for (uint32_t e = blockIdx.x; e < n; e += gridDim.x, par ^= 1u) {
        HitList& hits = hitl[par];
        // the other list, last read before the previous entry's end barrier, is made ready for the
        // next entry (which reads it after this entry's end barrier)
        if (tid == 0) {
            hitl[par ^ 1u].n = 0;
            hitl[par ^ 1u].ovf = 0;
            hitl[par ^ 1u].best = ~0ull;
        }
        BsgsEntry<F> be;
        be.E = pb.ent + (uint64_t)e * kEntWords;
        const uint64_t* __restrict__ E = be.E;
        // the entry's uniform words through the constant address space: scalar loads (k_setup wrote
        // the entries in an earlier launch; nothing writes them during this one)
        typedef __attribute__((address_space(4))) const uint64_t cu64;
        const cu64* Ec = reinterpret_cast<const cu64*>(reinterpret_cast<uintptr_t>(E));
        const uint64_t p = Ec[E_P];
        be.f.set(p, Ec[E_AUX], Ec[E_ONE]);
        const uint64_t mg = Ec[E_MG];
        be.m = (uint32_t)mg;
        be.ng = (uint32_t)(mg >> 32);
        const uint64_t ssrc = Ec[E_SSRC];
        be.s = (uint32_t)ssrc & 0xFFu;
        be.meta = (uint32_t)Ec[E_META];
        be.qp = be.meta & 0xFFu;
        be.nb = Ec[E_NB];
        be.klim = Ec[E_KLIM];
        be.t = F::from_bits(Ec[E_T]);
        be.B64 = F::from_bits(Ec[E_B64]);
        const F& f = be.f;

        // the entry uses as many buckets as its m needs, at least 16 and at most the launch's table:
        // the largest nbe with load * 2^(nbe-1) < m, i.e. bitlen(m - 1) - log2(load) (load is 2 or 4)
        const int nbe_m = (int)(32u - __clz(be.m - 1u)) - (int)(31u - __clz(tg.load));
        const uint32_t nbe = (uint32_t)max(4, min((int)tg.nbbits, nbe_m));
        const uint32_t nbk_e = 1u << nbe;
        Tb.omask = (nbk_e - 1u) << 4;
        const uint32_t spare = 20u * nbk + 2u * (uint32_t)sizeof(HitList);  // a halfword past the table

        // baby steps t * B^i, i == tid (mod L); a baby value equal to t for some 0 < i < m means
        // the order of B is below m: the giant steps might not reach the smallest solution, and
        // the host decides the prime.  Each lane counts its baby values equal to t; lane 0's first
        // one (i = 0) is t itself, so the order is below m when a lane counts more than that.
        uint32_t nt = 0;
        // the workgroup's next entry: one load per cache line of its first 48 words (the header, the
        // start tables and the strides) sets those lines on their way to the cache while this
        // entry's steps run; consumed at the end of the iteration, when it has long arrived
        uint32_t pf;
        {
            T bstep;
            T v = thread_pow<F, WB>(f, E, E_BLO, E_BHI, be.B64, E_BL, tid, bstep);  // t * B^tid
            __builtin_amdgcn_sched_barrier(0);  // issued after this entry's table loads
            {  // no branch around the load, so the wait for this entry's tables does not include it
                const uint32_t enx = e + gridDim.x < n ? e + gridDim.x : e;
                // one 32-bit word: a register the load writes stays live until the wait at the end
                pf = reinterpret_cast<const uint32_t*>(pb.ent)[((uint64_t)enx * kEntWords + min((tid & 3u) * 16u, 47u)) * 2];
            }
            __builtin_amdgcn_sched_barrier(0);
            // rounds in which every lane steps run on a uniform counter; lanes below the remainder
            // take one step more
            const uint32_t rounds = be.m / L, rem = be.m - rounds * L;
            const uint32_t wt = f.hash_word(be.t);
            auto baby = [&]() {  // baby step i = tid + r * L
                const uint32_t w = f.hash_word(v);
                uint32_t fp, bo;
                hash_word(w, Tb.omask, fp, bo);
                // one try in the key's own bucket, written to a spare halfword when the bucket is
                // full (no branch around the write); a full bucket (rare) sends the key on along the
                // chain
                const uint32_t old = atomicAdd(reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(Tb.cnt) + (bo >> 2)), 1u);
                // a baby value with t's hash word (equal words: equal low words of the residues);
                // one other than lane 0's first sends the prime to the host, which is exact, so a
                // word match of two different values costs time only
                nt += w == wt;
                v = f.mul(v, bstep);  // the next baby value while the atomic is in flight
                asm volatile("" : "+v"(v));  // (kept here: the compiler would sink it past the store)
                const bool placed = old < 8;
                *reinterpret_cast<lds_u16*>((size_t)(placed ? bo + 2u * old : spare)) = (uint16_t)fp;
                if (!placed) {
                    hits.ovf = 1u;  // every writer stores the same value
                    tab_insert(Tb, (bo + 16u) & Tb.omask, fp);
                }
            };
            for (uint32_t r = 0; r < rounds; ++r) baby();
            if (tid < rem) baby();
        }
        if (nt > (tid == 0 ? 1u : 0u)) hits.n = kCycleFlag;  // every writer stores the same value
        __syncthreads();
        if (hits.n >= kCycleFlag) {  // the same verdict in every lane
            if (tid == 0) put_record(pb, atomicAdd(&pb.cnt->n_out, 1u), pb.out_index ? (ssrc >> 32) : p, kHostDecides);
        } else {
            // giant steps G_j * (B^m)^g, g == tid (mod L), g < ng
            T mstep;
            const T Mt = thread_pow<F, WB>(f, E, E_MLO, E_MHI, F::from_bits(Ec[E_BM64]), E_BML, tid, mstep);
            const bool gfold = (be.meta >> 12) & 1u;  // the giant tables hold G * (B^m)^k
            const uint32_t rounds = be.ng / L, rem = be.ng - rounds * L;
            // with F::kTightLookup, a table in which no baby step overflowed its bucket (most entries:
            // at two keys per bucket about 6% of the tables of 512 keys have one) is looked up in the
            // key's own bucket only, so a lane walks the chain only on a fingerprint match, not for
            // every full bucket (a full bucket sends about one 2-step group of wavefront lookups in 8
            // down the walk at that load)
            auto giants = [&](auto chained) {
                for (uint32_t j = 0; j < be.s; ++j) {
                    T v = gfold ? Mt : f.mul(F::from_bits(Ec[E_G + j]), Mt);
                    auto probe = [&](uint32_t& hw, uint32_t& bo, T& vg) {
                        hw = f.hash_word(v);
                        bo = (hw >> 16) & Tb.omask;
                        const uint4 w = read_bucket(Tb, bo);
                        vg = v;
                        v = f.mul(v, mstep);  // the next giant value while the LDS read is in flight
                        if constexpr (decltype(chained)::value)
                            return bucket_needs_look(w, fp_pair(hw));
                        else
                            return bucket_may_hold(w, fp_pair(hw));
                    };
                    auto look = [&](uint32_t hw, uint32_t bo, T vg, uint32_t g) {
                        collect_hit(Tb, bo, min(hw & 0xFFFFu, 0xFFFEu), f.key(vg), g, j, hits);
                    };
                    uint32_t r = 0;
                    constexpr uint32_t G = F::kGiantGroup;
                    if constexpr (G > 1) {
                        for (; r + G <= rounds; r += G) {  // G steps, then one branch for all of them
                            uint32_t h[G], b[G];
                            T vs[G];
                            bool l[G], any = false;
#pragma unroll
                            for (uint32_t k = 0; k < G; ++k) {
                                l[k] = probe(h[k], b[k], vs[k]);
                                any |= l[k];
                            }
                            if (any) {
#pragma unroll
                                for (uint32_t k = 0; k < G; ++k)
                                    if (l[k]) look(h[k], b[k], vs[k], tid + (r + k) * L);
                            }
                        }
                    }
                    for (; r < rounds; ++r) {
                        uint32_t h, b;
                        T vg;
                        if (probe(h, b, vg)) look(h, b, vg, tid + r * L);
                    }
                    if (tid < rem) {
                        uint32_t h, b;
                        T vg;
                        if (probe(h, b, vg)) look(h, b, vg, tid + rounds * L);
                    }
                }
            };
            if (F::kTightLookup && hits.ovf == 0)  // the same in every lane (read after the babies' barrier)
                giants(std::integral_constant<bool, !F::kTightLookup>{});
            else
                giants(std::true_type{});
            __syncthreads();
            // exact check of every match: the solution k = g*m - i, when in range, gives
            // n = nj + (k*stride + parity)*cm; the smallest over the workgroup by an LDS minimum
            const uint32_t nh = hits.n;
            if (nh > kMaxHits) {
                if (tid == 0) put_record(pb, atomicAdd(&pb.cnt->n_out, 1u), pb.out_index ? (ssrc >> 32) : p, kHostDecides);
            } else {
                const bool qmode = (be.meta >> 10) & 1u;
                const uint64_t stride = qmode ? 1 : ((be.meta >> 8) & 1u) + 1;
                const uint64_t parity = qmode ? 0 : (be.meta >> 9) & 1u;
                const uint64_t cmul = qmode ? be.qp : ((be.meta >> 11) & 1u) ? Ec[E_DJ] : item_cm;
                // the baby steps again: every baby value t * B^i equal to a hit's giant value gives
                // the solution k = g*m - i
                T bstep;
                T v = thread_pow<F, WB>(f, E, E_BLO, E_BHI, be.B64, E_BL, tid, bstep);  // t * B^tid
                for (uint32_t i = tid; nh != 0 && i < be.m; i += L) {
                    const uint64_t x = f.key(v);
                    for (uint32_t h = 0; h < nh; ++h) {
                        const uint4 c = hits.h[h];
                        if (x != ((uint64_t)c.x | (uint64_t)c.y << 32)) continue;
                        const uint32_t j = c.w;
                        uint64_t nj, Kj;
                        if (qmode) {
                            const uint32_t dj = (uint32_t)(E[E_DJ + (j >> 3)] >> (8 * (j & 7))) & 0xFFu;
                            nj = be.nb + dj;
                            Kj = (item_N1 - nj + be.qp - 1) / be.qp;
                        } else {
                            nj = be.nb;
                            Kj = be.klim;
                        }
                        const int64_t kk = (int64_t)c.z * be.m - (int64_t)i;
                        if (kk < 0 || (uint64_t)kk >= Kj) continue;
                        atomicMin(&hits.best, (unsigned long long)(nj + ((uint64_t)kk * stride + parity) * cmul));
                    }
                    v = f.mul(v, bstep);
                }
                __syncthreads();
                // primality of the number is decided by k_prime_records, which runs next
                if (tid == 0 && hits.best != ~0ull)
                    put_record(pb, atomicAdd(&pb.cnt->n_out, 1u), pb.out_index ? (ssrc >> 32) : p, hits.best);
            }
        }
        // the buckets this entry used are emptied (the giant steps, the last readers of the table,
        // ended at a barrier)
        if (nbk_e >= L) {  // both powers of two: every lane clears nbk_e / L buckets
            for (uint32_t k = 0; k < nbk_e / L; ++k) {
                fp4[tid + k * L] = make_uint4(~0u, ~0u, ~0u, ~0u);
                Tb.cnt[tid + k * L] = 0u;
            }
        } else if (tid < nbk_e) {
            fp4[tid] = make_uint4(~0u, ~0u, ~0u, ~0u);
            Tb.cnt[tid] = 0u;
        }
        asm volatile("" ::"v"(pf));
        __syncthreads();
    }
}

// Deterministic Miller-Rabin with the 12 bases 2..37 (exact for n < 2^64, is_prime64
// Math.hpp:209-251), in Montgomery arithmetic; n odd, n < 2^62.
// One round of the deterministic 12-base Miller-Rabin test (bases 2..37, exact for n < 2^64,
// is_prime64 in Math.hpp): false when base a proves the odd number n > 37 composite (a divides n,
// or a is a strong witness).
__device__ __forceinline__ bool mr_round(uint64_t n, uint32_t a) {
    if (n % a == 0) return false;
    MontField f;
    f.init(n);
    const uint64_t one = f.one(), mone = n - one;  // Montgomery forms of 1 and -1
    uint64_t d = n - 1;
    const int s = __builtin_ctzll(d);
    d >>= s;
    uint64_t x = fpow(f, f.from(a, f.r2()), d);
    if (x == one || x == mone) return true;
    for (int r = 1; r < s; ++r) {
        x = f.mul(x, x);
        if (x == mone) return true;
    }
    return false;
}

// The relation records of a pass through the 12-base Miller-Rabin test: a record whose number is
// composite (a Fermat pseudoprime with a BSGS solution) gets number 0, which every reader of the
// records drops.  Records the host decides (which include every number below 1000), and records
// that carry an index instead of the number, stay as they are.  A kernel of its own so that k_bsgs
// does not carry the test's 64-bit arithmetic in its register budget; 16 lanes per record run its
// 12 rounds side by side (a round is a chain of dependent products: one lane per round keeps the
// kernel short).
__global__ __launch_bounds__(256) void k_prime_records(PassBuffers pb) {
    if (pb.out_index) return;
    const uint32_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    const uint32_t n = min(pb.cnt->n_out, pb.out_cap);
    const uint32_t k = threadIdx.x & 15u, grp = (threadIdx.x >> 4) & 3u;
    for (uint32_t r0 = blockIdx.x * 16; r0 < n; r0 += gridDim.x * 16) {  // uniform over the block
        const uint32_t i = r0 + (threadIdx.x >> 4);
        const bool live = i < n && pb.out[2 * (uint64_t)i + 1] != kHostDecides;
        const bool composite = live && k < 12 && !mr_round(pb.out[2 * (uint64_t)i], bases[k]);
        const uint64_t votes = __ballot(composite);
        if (live && k == 0 && ((votes >> (16 * grp)) & 0xFFFFull) != 0) pb.out[2 * (uint64_t)i] = 0;
    }
}

void launch_prime_records(hipStream_t s, const PassBuffers& pb) {
    hipLaunchKernelGGL(k_prime_records, dim3(kPrimeRecordBlocks), dim3(256), 0, s, pb);
}

// ------------------------------------------------------------------------------------- launchers

void launch_reset(hipStream_t s, Counters* cnt, uint32_t n_surv0) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, cnt, n_surv0);
}

void launch_sieve(hipStream_t s, const PassDesc* d, uint64_t nodd, const SieveTable& tab, const PassBuffers& pb,
                  uint32_t grid_override) {
    const uint32_t grid = grid_override ? grid_override : (uint32_t)((nodd + kSegOdds - 1) / kSegOdds);
    hipLaunchKernelGGL(k_sieve, dim3(grid), dim3(kSieveThreads), 0, s, d, tab, pb);
}

void launch_filter(hipStream_t s, const PassDesc* d, const PassBuffers& pb, uint32_t grid) {
    hipLaunchKernelGGL(k_filter, dim3(grid), dim3(256), 0, s, d, pb);
}

template <class F>
const void* setup_kernel_f(uint32_t mode) {
    return mode == kModeOrderExt ? reinterpret_cast<const void*>(&k_setup_order<F, true>)
           : mode == kModeOrder ? reinterpret_cast<const void*>(&k_setup_order<F, false>)
           : mode == kModeClass ? reinterpret_cast<const void*>(&k_setup<F, kModeClass>)
                                : reinterpret_cast<const void*>(&k_setup<F, kModePlain>);
}

const void* setup_kernel(FieldKind fk, uint32_t mode) {
    return fk == FIELD_F64 ? setup_kernel_f<F64Field>(mode) : setup_kernel_f<MontField>(mode);
}

void launch_setup(hipStream_t s, FieldKind fk, uint32_t mode, const PassDesc* d, const PassBuffers& pb, uint32_t grid) {
    PassBuffers b = pb;
    void* args[] = {&d, &b};
    hipLaunchKernel(setup_kernel(fk, mode), dim3(grid), dim3(256), args, 0, s);
}

// Table geometry for baby-step counts up to m_max: at most `load` keys per 8-slot bucket on
// average (2: a quarter of the slots filled, 4: half).
TableGeom table_geom(uint64_t m_max, uint64_t waves, uint32_t load) {
    TableGeom tg;
    tg.ibits = 16;
    tg.fpbits = 16;
    tg.nbbits = 4;
    while (((uint64_t)load << tg.nbbits) < m_max) ++tg.nbbits;
    tg.waves = (uint32_t)waves;
    tg.load = load;
    return tg;
}

// fingerprints (16 B) + a fill counter (4 B) per bucket, two hit lists, and 16 bytes that hold the
// spare halfword of the baby steps' inserts
size_t bsgs_lds_bytes(const TableGeom& tg) { return ((size_t)20 << tg.nbbits) + 2 * sizeof(HitList) + 16; }

// n_ent: entries, one workgroup each (0 = a persistent grid of `grid` workgroups that drains the
// pass's entries).
template <class F>
void launch_bsgs_f(hipStream_t s, const TableGeom& tg, const PassDesc* d, const PassBuffers& pb, uint32_t grid,
                   size_t lds) {
    if (tg.waves == 1)
        hipLaunchKernelGGL((k_bsgs<F, 1>), dim3(grid), dim3(64), lds, s, d, pb);
    else if (tg.waves == 2)
        hipLaunchKernelGGL((k_bsgs<F, 2>), dim3(grid), dim3(128), lds, s, d, pb);
    else
        hipLaunchKernelGGL((k_bsgs<F, 4>), dim3(grid), dim3(256), lds, s, d, pb);
}

template <class F>
int bsgs_blocks_per_cu(const TableGeom& tg, size_t lds) {
    int nb = 0;
    hipError_t e;
    if (tg.waves == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bsgs<F, 1>, 64, lds);
    else if (tg.waves == 2)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bsgs<F, 2>, 128, lds);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bsgs<F, 4>, 256, lds);
    return e == hipSuccess ? nb : 0;
}

uint32_t bsgs_persistent_grid(FieldKind fk, const TableGeom& tg, int cus) {
    const size_t lds = bsgs_lds_bytes(tg);
    int nb = fk == FIELD_F64 ? bsgs_blocks_per_cu<F64Field>(tg, lds) : bsgs_blocks_per_cu<MontField>(tg, lds);
    if (nb <= 0) nb = 1;
    // kBsgsOversubscribe times the resident workgroups: each gets a few entries, and the hardware
    // starts the next workgroup as one finishes, so a workgroup with costly entries does not hold
    // up the pass (with exactly the resident count, the slowest static share set the pass time)
    return kBsgsOversubscribe * (uint32_t)nb * (uint32_t)std::max(1, cus);
}

bool bsgs_lds_at_zero() {
    const void* k[6] = {reinterpret_cast<const void*>(&k_bsgs<F64Field, 1>), reinterpret_cast<const void*>(&k_bsgs<F64Field, 2>),
                        reinterpret_cast<const void*>(&k_bsgs<F64Field, 4>), reinterpret_cast<const void*>(&k_bsgs<MontField, 1>),
                        reinterpret_cast<const void*>(&k_bsgs<MontField, 2>), reinterpret_cast<const void*>(&k_bsgs<MontField, 4>)};
    for (const void* f : k) {
        hipFuncAttributes a;
        if (hipFuncGetAttributes(&a, f) != hipSuccess || a.sharedSizeBytes != 0) return false;
    }
    return true;
}

// The BSGS kernel of a variant and its workgroup size.
const void* bsgs_kernel(FieldKind fk, uint32_t waves, uint32_t& threads) {
    threads = 64 * (waves == 1 ? 1 : waves == 2 ? 2 : 4);
    if (fk == FIELD_F64)
        return waves == 1   ? reinterpret_cast<const void*>(&k_bsgs<F64Field, 1>)
               : waves == 2 ? reinterpret_cast<const void*>(&k_bsgs<F64Field, 2>)
                            : reinterpret_cast<const void*>(&k_bsgs<F64Field, 4>);
    return waves == 1   ? reinterpret_cast<const void*>(&k_bsgs<MontField, 1>)
           : waves == 2 ? reinterpret_cast<const void*>(&k_bsgs<MontField, 2>)
                        : reinterpret_cast<const void*>(&k_bsgs<MontField, 4>);
}


hipError_t enqueue_pass_chain(hipStream_t s, const PassChain& c) {
    hipError_t e = hipMemcpyAsync(c.d_desc, c.h_desc, sizeof(PassDesc), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    launch_reset(s, c.pb.cnt, 0);
    launch_sieve(s, c.d_desc, 0, c.tab, c.pb, c.grid_sieve);
    launch_filter(s, c.d_desc, c.pb, c.grid_stage);
    launch_setup(s, c.fk, c.mode, c.d_desc, c.pb, c.grid_stage);
    launch_bsgs(s, c.fk, c.tg, c.d_desc, c.pb, 0, c.grid_bsgs);
    launch_prime_records(s, c.pb);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(c.h_cnt, c.pb.cnt, sizeof(Counters), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(c.h_out, c.pb.out, c.out_words * sizeof(uint64_t), hipMemcpyDeviceToHost, s);
}

namespace {

hipError_t add_kernel(hipGraph_t g, hipGraphNode_t& prev, const void* fn, uint32_t grid, uint32_t threads,
                      uint32_t lds, void** args) {
    hipKernelNodeParams kp{};
    kp.func = const_cast<void*>(fn);
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(threads);
    kp.sharedMemBytes = lds;
    kp.kernelParams = args;
    kp.extra = nullptr;
    hipGraphNode_t node;
    const hipError_t e = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
    if (e == hipSuccess) prev = node;
    return e;
}

hipError_t add_copy(hipGraph_t g, hipGraphNode_t& prev, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipGraphNode_t node;
    const hipError_t e = hipGraphAddMemcpyNode1D(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, dst, src, bytes, kind);
    if (e == hipSuccess) prev = node;
    return e;
}

}  // namespace

hipError_t build_pass_graph(const PassChain& c, hipGraph_t* out) {
    hipGraph_t g = nullptr;
    hipError_t e = hipGraphCreate(&g, 0);
    if (e != hipSuccess) return e;
    hipGraphNode_t prev = nullptr;
    // kernel arguments by address, in declaration order
    const PassDesc* d = c.d_desc;
    Counters* cnt = c.pb.cnt;
    uint32_t zero = 0;
    PassBuffers pb = c.pb;
    SieveTable tab = c.tab;
    void* a_reset[] = {&cnt, &zero};
    void* a_sieve[] = {&d, &tab, &pb};
    void* a_stage[] = {&d, &pb};
    void* a_records[] = {&pb};
    uint32_t bsgs_threads = 0;
    const void* bsgs = bsgs_kernel(c.fk, c.tg.waves, bsgs_threads);
    if (e == hipSuccess) e = add_copy(g, prev, c.d_desc, c.h_desc, sizeof(PassDesc), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = add_kernel(g, prev, reinterpret_cast<const void*>(&k_reset), 1, 64, 0, a_reset);
    if (e == hipSuccess)
        e = add_kernel(g, prev, reinterpret_cast<const void*>(&k_sieve), c.grid_sieve, kSieveThreads, 0, a_sieve);
    if (e == hipSuccess) e = add_kernel(g, prev, reinterpret_cast<const void*>(&k_filter), c.grid_stage, 256, 0, a_stage);
    if (e == hipSuccess) e = add_kernel(g, prev, setup_kernel(c.fk, c.mode), c.grid_stage, 256, 0, a_stage);
    if (e == hipSuccess)
        e = add_kernel(g, prev, bsgs, c.grid_bsgs, bsgs_threads, (uint32_t)bsgs_lds_bytes(c.tg), a_stage);
    if (e == hipSuccess)
        e = add_kernel(g, prev, reinterpret_cast<const void*>(&k_prime_records), kPrimeRecordBlocks, 256, 0, a_records);
    if (e == hipSuccess) e = add_copy(g, prev, c.h_cnt, c.pb.cnt, sizeof(Counters), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = add_copy(g, prev, c.h_out, c.pb.out, c.out_words * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return e;
    }
    *out = g;
    return hipSuccess;
}

void launch_bsgs(hipStream_t s, FieldKind fk, const TableGeom& tg, const PassDesc* d, const PassBuffers& pb,
                 uint32_t n_ent, uint32_t grid) {
    const size_t lds = bsgs_lds_bytes(tg);
    if (n_ent) grid = n_ent;
    if (grid == 0) return;
    if (fk == FIELD_F64)
        launch_bsgs_f<F64Field>(s, tg, d, pb, grid, lds);
    else
        launch_bsgs_f<MontField>(s, tg, d, pb, grid, lds);
}

}  // namespace mps
