// HIP kernels of the MI355X sieve worker (gfx950, wave64).
//
// Stages, one pass = one range of odd numbers [lo, lo + 2*nodd):
//   k_sieve   segmented odd sieve, one workgroup per 2^17-odd LDS bitmap; odd primes 3..65521 cross
//             off their odd multiples from max(q^2, first multiple in the segment), so a prime is
//             never crossed off; survivors are compacted by a block scan of popcounts.
//   k_filter  per survivor: Jacobi-symbol residue filter (plain mode) or per-prime class split with
//             a quadratic-residue mask (class mode); wave ballot compaction of what is left.
//   k_setup   per candidate: base-2 Fermat test, g-th power residue test, BSGS sizes and the entry
//             (powers tables, giant starts), compacted by wave ballot.
//   k_bsgs    one wavefront per entry: baby steps t*B^i into an LDS fingerprint table, giant steps
//             G_j*(B^m)^g looked up; a hit appends p to the flagged list.
// Numbers the device cannot decide (p < 1000, p | b) are appended to the flagged list directly.
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

// Index j of the first odd multiple v0 + 2j of q that is >= max(v0, q^2); v0 odd.
__device__ __forceinline__ uint32_t seg_start(uint64_t v0, uint32_t q, uint32_t r32) {
    const uint32_t r = mod_small(v0, q, r32);
    uint32_t j = ((q - r) % q) * ((q + 1) >> 1) % q;  // 2j == -v0 (mod q)
    const uint64_t qq = (uint64_t)q * q;
    if (v0 + 2 * (uint64_t)j < qq) {
        const uint64_t jj = (qq - v0) >> 1;
        j = jj > 0x7FFFFFFFull ? 0x7FFFFFFFu : (uint32_t)jj;
    }
    return j;
}

__global__ __launch_bounds__(kSieveThreads) void k_sieve(uint64_t lo, uint64_t nodd, const uint32_t* __restrict__ sp,
                                                         const uint32_t* __restrict__ sp_r32, uint32_t nsp,
                                                         uint32_t nsmall, PassBuffers pb) {
    __shared__ uint32_t bits[kSegOdds / 32];
    __shared__ uint32_t st[kSmallPrimeBound / 2];
    __shared__ uint32_t wtot[kSieveThreads / 64];
    __shared__ uint32_t gbase;
    const uint32_t tid = threadIdx.x;
    const uint64_t seg0 = (uint64_t)blockIdx.x * kSegOdds;
    if (seg0 >= nodd) return;
    const uint64_t rest = nodd - seg0;
    const uint32_t len = rest < kSegOdds ? (uint32_t)rest : kSegOdds;
    const uint64_t v0 = lo + 2 * seg0;

    uint4* b4 = reinterpret_cast<uint4*>(bits);
    for (uint32_t i = tid; i < kSegOdds / 128; i += kSieveThreads) b4[i] = make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t k = tid; k < nsmall; k += kSieveThreads) st[k] = seg_start(v0, sp[k], sp_r32[k]);
    __syncthreads();

    // primes below kSmallPrimeBound: the lanes of the block share each prime
    for (uint32_t k = 0; k < nsmall; ++k) {
        const uint32_t q = sp[k];
        for (uint32_t i = st[k] + tid * q; i < len; i += kSieveThreads * q) atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    // larger primes: one lane per prime
    for (uint32_t k = nsmall + tid; k < nsp; k += kSieveThreads) {
        const uint32_t q = sp[k];
        for (uint32_t i = seg_start(v0, q, sp_r32[k]); i < len; i += q) atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    __syncthreads();

    // compaction: lane t owns words t, t + 256, ...
    const uint32_t words = (len + 31) >> 5;
    uint32_t mine = 0;
    for (uint32_t w = tid; w < words; w += kSieveThreads) {
        uint32_t x = ~bits[w];
        const uint32_t valid = len - (w << 5);
        if (valid < 32) x &= (1u << valid) - 1u;
        mine += __popc(x);
    }
    const uint32_t lane = lane_id(), wave = tid >> 6;
    uint32_t incl = mine;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    if (tid == 0) {
        uint32_t tot = 0;
        for (uint32_t i = 0; i < kSieveThreads / 64; ++i) tot += wtot[i];
        gbase = atomicAdd(&pb.cnt->n_surv, tot);
    }
    __syncthreads();
    uint32_t pos = gbase + incl - mine;
    for (uint32_t i = 0; i < wave; ++i) pos += wtot[i];
    for (uint32_t w = tid; w < words; w += kSieveThreads) {
        uint32_t x = ~bits[w];
        const uint32_t valid = len - (w << 5);
        if (valid < 32) x &= (1u << valid) - 1u;
        while (x) {
            const uint32_t b = __builtin_ctz(x);
            if (pos < pb.surv_cap) pb.surv[pos] = (uint32_t)(seg0 + (w << 5) + b);
            ++pos;
            x &= x - 1u;
        }
    }
}

// ------------------------------------------------------------------------------------- filter

__device__ __forceinline__ uint64_t mod_signed(int64_t x, uint64_t p) {  // x mod p in [0, p)
    if (x >= 0) {
        const uint64_t u = (uint64_t)x;
        return u < p ? u : u % p;
    }
    const uint64_t a = (uint64_t)(-x);
    const uint64_t r = a < p ? a : a % p;
    return r == 0 ? 0 : p - r;
}

// (p - 1) mod d for the class modulus: 32-bit path when d < 2^16.
__device__ __forceinline__ uint64_t pm1_mod(uint64_t p, uint64_t d, uint32_t r32) {
    if (d < 65536u) return mod_small(p - 1, (uint32_t)d, r32);
    return (p - 1) % d;
}

__device__ __forceinline__ uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// Plain mode (prep_core decision, Gpu.hpp:231-263).  info: bit0 = stride 2, bit1 = odd k.
__device__ __forceinline__ void filter_plain(uint64_t p, const ItemParams& ip, bool& keep, bool& flag2,
                                             uint32_t& info) {
    if (p < 1000 || (p & 1) == 0) { flag2 = true; return; }
    if (mod_signed(ip.b, p) == 0) { flag2 = true; return; }
    if (mod_signed(ip.c, p) == 0) return;  // t = -c == 0: b^n == 0 impossible
    if (ip.kcount == 0) return;
    const uint64_t g = gcd64(ip.cm, pm1_mod(p, ip.cm, ip.cm_r32));
    const int jb = jacobi_signed(ip.b, p);
    const int jt = jacobi_signed(-ip.c, p);
    const int lT = jt * ((ip.n_start & 1) ? jb : 1);
    if (g == 2) {
        keep = lT == 1;
    } else if (g < 2) {
        if (jb == 1) {
            keep = lT == 1;
        } else {
            keep = true;
            info = 1u | (lT != 1 ? 2u : 0u);
        }
    } else {
        keep = true;
    }
}

struct ClassSplit {
    uint32_t g, qp, ncls;
    uint32_t dj[kMaxClasses];
};

__device__ __forceinline__ void class_split(uint64_t p, const ItemParams& ip, ClassSplit& cs) {
    const uint32_t Q = (uint32_t)ip.Qmax;
    const uint32_t cm = (uint32_t)ip.cm;
    cs.g = gcd32(Q, mod_small(p - 1, Q, ip.Q_r32));
    cs.qp = cs.g / gcd32(cs.g, cm) * cm;
    cs.ncls = cs.qp / cm;
    const uint32_t n0q = ip.n0modQ % cs.qp;
    const uint32_t c0 = (uint32_t)(ip.cres % cm);
    for (uint32_t j = 0; j < cs.ncls; ++j) cs.dj[j] = (c0 + j * cm + cs.qp - n0q) % cs.qp;
}

// Class mode (prep_q_core decision, Gpu.hpp:478-509).  info = class mask.
__device__ __forceinline__ void filter_class(uint64_t p, const ItemParams& ip, bool& keep, bool& flag2,
                                             uint32_t& info) {
    if (p < 1000 || (p & 1) == 0) { flag2 = true; return; }
    if (mod_signed(ip.b, p) == 0) { flag2 = true; return; }
    if (mod_signed(ip.c, p) == 0) return;
    ClassSplit cs;
    class_split(p, ip, cs);
    const uint64_t span = ip.N1 - ip.N0;
    const int jb = jacobi_signed(ip.b, p);
    const int jt = jacobi_signed(-ip.c, p);
    const uint32_t n0odd = (uint32_t)(ip.N0 & 1);
    uint32_t mask = 0;
    for (uint32_t j = 0; j < cs.ncls; ++j) {
        if ((uint64_t)cs.dj[j] >= span) continue;
        const int l = jt * (((n0odd ^ (cs.dj[j] & 1u)) != 0) ? jb : 1);
        if (l == 1) mask |= 1u << j;
    }
    keep = mask != 0;
    info = mask;
}

__global__ __launch_bounds__(256) void k_filter(uint64_t lo, ItemParams ip, PassBuffers pb) {
    const uint32_t n = min(pb.cnt->n_surv, pb.surv_cap);
    const uint32_t lane = lane_id();
    const uint32_t wglobal = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i0 = wglobal * 64; i0 < n; i0 += nwaves * 64) {
        const uint32_t i = i0 + lane;
        bool keep = false, flag2 = false;
        uint32_t info = 0, idx = 0;
        uint64_t p = 0;
        if (i < n) {
            idx = pb.vals ? i : pb.surv[i];
            p = pb.vals ? pb.vals[i] : lo + 2 * (uint64_t)idx;
            if (ip.qmode)
                filter_class(p, ip, keep, flag2, info);
            else
                filter_plain(p, ip, keep, flag2, info);
        }
        const uint32_t slot = wave_append(keep, &pb.cnt->n_cand);
        if (keep && slot < pb.cand_cap) pb.cand[slot] = (uint64_t)idx | ((uint64_t)info << 32);
        const uint32_t oslot = wave_append(flag2, &pb.cnt->n_out);
        if (flag2 && oslot < pb.out_cap) pb.out[oslot] = pb.out_index ? idx : p;
        const uint64_t fb = __ballot(flag2);
        if (lane == 0 && fb != 0) atomicAdd(&pb.cnt->n_flag2, (uint32_t)__popcll(fb));
    }
}

// ------------------------------------------------------------------------------------- setup

__device__ __forceinline__ uint64_t isqrt_ceil64(uint64_t x) {
    if (x == 0) return 0;
    uint64_t r = (uint64_t)sqrt((double)x);
    while (r > 0 && r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return r * r < x ? r + 1 : r;
}

// Class mode sizes: Gc = 64*cr giant steps per class and m <= kClassMaxM baby steps with
// (Gc - 1) * m >= K - 1, chosen to minimise the wavefront rounds ceil(m/64) + s*cr.
__device__ __forceinline__ void class_sizes(uint64_t K, uint32_t s, uint64_t& m, uint64_t& Gc) {
    uint64_t crmin = (K + 64ull * (kClassMaxM - 1) - 1) / (64ull * (kClassMaxM - 1));
    if (crmin < 2) crmin = 2;
    const uint64_t crc = isqrt_ceil64(K / (4096ull * s));
    uint64_t best = ~0ull;
    m = kClassMaxM;
    Gc = 64 * crmin;
    const uint64_t c0 = crc > 2 ? crc - 2 : 1;
    for (uint64_t cr = c0; cr <= crc + 3; ++cr) {
        const uint64_t cx = cr < crmin ? crmin : cr;
        const uint64_t gc = 64 * cx;
        const uint64_t mm = (K + gc - 3) / (gc - 1) + 1;
        if (mm > kClassMaxM) continue;
        const uint64_t rounds = (mm + 63) / 64 + (uint64_t)s * cx;
        if (rounds < best) {
            best = rounds;
            m = mm;
            Gc = gc;
        }
    }
    if (best == ~0ull) {
        for (uint64_t cx = crmin;; ++cx) {
            const uint64_t gc = 64 * cx;
            const uint64_t mm = (K + gc - 3) / (gc - 1) + 1;
            if (mm <= kClassMaxM) {
                m = mm;
                Gc = gc;
                break;
            }
        }
    }
}

// Writes the powers tables of base x: tab[k] = x^k and tab[8 + k] = x^(8k), k < 8; returns x^64.
template <class F>
__device__ __forceinline__ typename F::T write_pow_tables(const F& f, typename F::T x, uint64_t* tab,
                                                         unsigned long long& work) {
    typedef typename F::T T;
    T a = f.one();
    for (int k = 0; k < 8; ++k) {
        tab[k] = F::to_bits(a);
        a = f.mul(a, x);
    }
    const T x8 = a;
    T c = f.one();
    for (int k = 0; k < 8; ++k) {
        tab[8 + k] = F::to_bits(c);
        c = f.mul(c, x8);
    }
    work += 16;
    return c;  // x^64
}

template <class F, bool QMODE>
__global__ __launch_bounds__(256) void k_setup(uint64_t lo, ItemParams ip, PassBuffers pb) {
    typedef typename F::T T;
    const uint32_t n = min(pb.cnt->n_cand, pb.cand_cap);
    const uint32_t lane = lane_id();
    const uint32_t wglobal = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    unsigned long long w_fermat = 0, w_setup = 0, w_bsgs = 0;
    uint32_t fermat_fail = 0;
    for (uint32_t i0 = wglobal * 64; i0 < n; i0 += nwaves * 64) {
        const uint32_t i = i0 + lane;
        bool keep = false;
        uint64_t p = 0, m = 0, ng = 0;
        uint32_t idx = 0, info = 0, s = 1, mask = 0, dmin = 0;
        ClassSplit cs;
        F f;
        if (i < n) {
            const uint64_t cd = pb.cand[i];
            idx = (uint32_t)cd;
            info = (uint32_t)(cd >> 32);
            p = pb.vals ? pb.vals[idx] : lo + 2 * (uint64_t)idx;
            f.init(p);
            if (pb.need_prime) w_fermat += 64 - __builtin_clzll(p - 1);
            if (pb.need_prime && !fermat2(f, p)) {
                ++fermat_fail;
            } else {
                const uint64_t bmod = mod_signed(ip.b, p), t = mod_signed(-ip.c, p);
                const T rr = f.r2();
                const T bm = f.from(bmod, rr), tm = f.from(t, rr);
                if (QMODE) {
                    class_split(p, ip, cs);
                    mask = info;
                    if (cs.g > 2) {
                        // class j needs t^e == (b^e)^(n_j mod g), e = (p-1)/g
                        const uint64_t e = (p - 1) / cs.g;
                        const T u = fpow(f, tm, e), w = fpow(f, bm, e);
                        w_setup += 2 * (64 - __builtin_clzll(e)) + cs.g;
                        const uint64_t ku = f.key(u);
                        const uint32_t n0g = ip.n0modQ % cs.g;
                        uint32_t keepm = 0;
                        T wk = f.one();
                        for (uint32_t k = 0; k < cs.g; ++k) {
                            if (f.key(wk) == ku)
                                for (uint32_t j = 0; j < cs.ncls; ++j)
                                    if ((n0g + cs.dj[j]) % cs.g == k) keepm |= 1u << j;
                            wk = f.mul(wk, w);
                        }
                        mask &= keepm;
                    }
                    if (mask != 0) {
                        s = __popc(mask);
                        dmin = cs.qp;
                        for (uint32_t j = 0; j < cs.ncls; ++j)
                            if (((mask >> j) & 1u) && cs.dj[j] < dmin) dmin = cs.dj[j];
                        const uint64_t K = (ip.N1 - ip.N0 + cs.qp - 1) / cs.qp;
                        class_sizes(K, s, m, ng);
                        keep = true;
                    }
                } else {
                    const uint64_t g = gcd64(ip.cm, pm1_mod(p, ip.cm, ip.cm_r32));
                    bool ok = true;
                    if (g > 2) {
                        // B = b^cm is a g-th power, so T = t*b^-n_start must be one: t^e == G0^e
                        const uint64_t e = (p - 1) / g;
                        const T G0 = fpow(f, bm, ip.n_start);
                        ok = f.key(fpow(f, tm, e)) == f.key(fpow(f, G0, e));
                        w_setup += 2 * (64 - __builtin_clzll(e)) + (64 - __builtin_clzll(ip.n_start | 1));
                    }
                    const uint32_t stride = (info & 1u) ? 2 : 1, parity = (info >> 1) & 1u;
                    uint64_t kc = ip.kcount < p ? ip.kcount : p;
                    if (ok && kc > parity) {
                        const uint64_t jcount = (kc - parity + stride - 1) / stride;
                        m = isqrt_ceil64(jcount);
                        if (m == 0) m = 1;
                        ng = (jcount + m - 1) / m + 1;  // g = 0 .. ceil(jcount/m)
                        keep = true;
                    }
                }
            }
        }
        const uint32_t slot = wave_append(keep, &pb.cnt->n_ent);
        if (keep && slot >= pb.ent_cap) {
            const uint32_t o = atomicAdd(&pb.cnt->n_out, 1u);  // no room: the host decides this one
            if (o < pb.out_cap) pb.out[o] = pb.out_index ? idx : p;
            keep = false;
        }
        if (keep) {
            uint64_t* E = pb.ent + (uint64_t)slot * kEntWords;
            const uint64_t bmod = mod_signed(ip.b, p), t = mod_signed(-ip.c, p);
            const T rr = f.r2();
            const T bm = f.from(bmod, rr), tm = f.from(t, rr);
            T B;
            E[E_P] = p;
            E[E_AUX] = f.aux();
            E[E_ONE] = F::to_bits(f.one());
            E[E_T] = F::to_bits(tm);
            E[E_MG] = m | (ng << 32);
            E[E_SSRC] = (uint64_t)s | ((uint64_t)((QMODE ? mask : info) & 0xFFFFFFu) << 8) | ((uint64_t)idx << 32);
            if (QMODE) {
                B = fpow(f, bm, cs.qp);
                // G_j = b^(N0 + dj) = b^(N0 + dmin) * b^(dj - dmin)
                const T base = fpow(f, bm, ip.N0 + dmin);
                w_setup += 2 * (64 - __builtin_clzll(ip.N0 + dmin + 1)) + 12;
                uint32_t o = 0;
                for (uint32_t j = 0; j < cs.ncls; ++j) {
                    if (!((mask >> j) & 1u)) continue;
                    E[E_G + o] = F::to_bits(f.mul(base, fpow(f, bm, cs.dj[j] - dmin)));
                    w_setup += 12;
                    ++o;
                }
            } else {
                B = fpow(f, bm, ip.cm);
                T G0 = fpow(f, bm, ip.n_start);
                w_setup += 2 * (64 - __builtin_clzll(ip.n_start | 1)) + 2 * (64 - __builtin_clzll(ip.cm));
                if (info & 1u) {  // stride 2: k = 2j + parity
                    if (info & 2u) G0 = f.mul(G0, B);
                    B = f.mul(B, B);
                }
                E[E_G] = F::to_bits(G0);
            }
            const T Bm = fpow(f, B, m);
            w_setup += 2 * (64 - __builtin_clzll(m));
            E[E_B64] = F::to_bits(write_pow_tables(f, B, E + E_BLO, w_setup));
            E[E_BM64] = F::to_bits(write_pow_tables(f, Bm, E + E_MLO, w_setup));
            w_bsgs += m + (uint64_t)s * ng;
        }
    }
    w_fermat = wave_sum64(w_fermat);
    w_setup = wave_sum64(w_setup);
    w_bsgs = wave_sum64(w_bsgs);
    const unsigned long long ff = wave_sum64(fermat_fail);
    if (lane == 0) {
        if (w_fermat) atomicAdd(&pb.cnt->w_fermat, w_fermat);
        if (w_setup) atomicAdd(&pb.cnt->w_setup, w_setup);
        if (w_bsgs) atomicAdd(&pb.cnt->w_bsgs, w_bsgs);
        if (ff) atomicAdd(&pb.cnt->n_fermat_fail, (uint32_t)ff);
    }
}

// ------------------------------------------------------------------------------------- bsgs

template <class F, int W>
__global__ __launch_bounds__(64 * W) void k_bsgs(PassBuffers pb, uint32_t tbits) {
    typedef typename F::T T;
    extern __shared__ uint32_t lds_tab[];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t tsize = 1u << tbits, tmask = tsize - 1u;
    uint32_t* tab = lds_tab + wv * tsize;
    const uint32_t n = min(pb.cnt->n_ent, pb.ent_cap);
    for (uint32_t e0 = blockIdx.x * W + wv; e0 < n; e0 += gridDim.x * W) {
        const uint32_t e = __builtin_amdgcn_readfirstlane(e0);
        const uint64_t* __restrict__ E = pb.ent + (uint64_t)e * kEntWords;
        const uint64_t p = E[E_P];
        F f;
        f.set(p, E[E_AUX], E[E_ONE]);
        const uint64_t mg = E[E_MG];
        const uint32_t m = (uint32_t)mg, ng = (uint32_t)(mg >> 32);
        const uint64_t ssrc = E[E_SSRC];
        const uint32_t s = (uint32_t)ssrc & 0xFFu;

        uint4* t4 = reinterpret_cast<uint4*>(tab);
        for (uint32_t i = lane; i < (tsize >> 2); i += 64) t4[i] = make_uint4(~0u, ~0u, ~0u, ~0u);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();

        // baby steps t * B^i, i = lane, lane + 64, ...
        const T B64 = F::from_bits(E[E_B64]);
        T v = f.mul(F::from_bits(E[E_T]),
                    f.mul(F::from_bits(E[E_BLO + (lane & 7)]), F::from_bits(E[E_BHI + (lane >> 3)])));
        for (uint32_t i = lane; i < m; i += 64) {
            const uint32_t w = table_word(f.key(v));
            uint32_t h = w & tmask;
            for (;;) {
                const uint32_t prev = atomicCAS(&tab[h], 0xFFFFFFFFu, w);
                if (prev == 0xFFFFFFFFu || prev == w) break;
                h = (h + 1) & tmask;
            }
            v = f.mul(v, B64);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();

        // giant steps G_j * (B^m)^g, g = lane, lane + 64, ... < ng
        const T M64 = F::from_bits(E[E_BM64]);
        const T Ml = f.mul(F::from_bits(E[E_MLO + (lane & 7)]), F::from_bits(E[E_MHI + (lane >> 3)]));
        bool found = false;
        for (uint32_t j = 0; j < s; ++j) {
            v = f.mul(F::from_bits(E[E_G + j]), Ml);
            for (uint32_t g = lane; g < ng; g += 64) {
                const uint32_t w = table_word(f.key(v));
                uint32_t h = w & tmask;
                for (;;) {
                    const uint32_t x = tab[h];
                    if (x == w) {
                        found = true;
                        break;
                    }
                    if (x == 0xFFFFFFFFu) break;
                    h = (h + 1) & tmask;
                }
                v = f.mul(v, M64);
            }
        }
        if (__ballot(found) != 0 && lane == 0) {
            const uint32_t o = atomicAdd(&pb.cnt->n_out, 1u);
            if (o < pb.out_cap) pb.out[o] = pb.out_index ? (ssrc >> 32) : p;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------------------------------- launchers

void launch_reset(hipStream_t s, Counters* cnt, uint32_t n_surv0) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, cnt, n_surv0);
}

void launch_sieve(hipStream_t s, uint64_t lo, uint64_t nodd, const uint32_t* sp, const uint32_t* sp_r32,
                  uint32_t nsp, uint32_t nsmall, const PassBuffers& pb, uint32_t grid_override) {
    const uint32_t grid = grid_override ? grid_override : (uint32_t)((nodd + kSegOdds - 1) / kSegOdds);
    hipLaunchKernelGGL(k_sieve, dim3(grid), dim3(kSieveThreads), 0, s, lo, nodd, sp, sp_r32, nsp, nsmall, pb);
}

void launch_filter(hipStream_t s, uint64_t lo, const ItemParams& ip, const PassBuffers& pb, uint32_t grid) {
    hipLaunchKernelGGL(k_filter, dim3(grid), dim3(256), 0, s, lo, ip, pb);
}

void launch_setup(hipStream_t s, FieldKind fk, uint64_t lo, const ItemParams& ip, const PassBuffers& pb,
                  uint32_t grid) {
    if (fk == FIELD_F64) {
        if (ip.qmode)
            hipLaunchKernelGGL((k_setup<F64Field, true>), dim3(grid), dim3(256), 0, s, lo, ip, pb);
        else
            hipLaunchKernelGGL((k_setup<F64Field, false>), dim3(grid), dim3(256), 0, s, lo, ip, pb);
    } else {
        if (ip.qmode)
            hipLaunchKernelGGL((k_setup<MontField, true>), dim3(grid), dim3(256), 0, s, lo, ip, pb);
        else
            hipLaunchKernelGGL((k_setup<MontField, false>), dim3(grid), dim3(256), 0, s, lo, ip, pb);
    }
}

// n_ent: entries to cover with one wavefront each (0 = a persistent grid of `grid` blocks).
// Two wavefronts (two tables) per block up to 2^12-word tables, one above, so that a block never
// asks for more than 64 KiB of LDS.
void launch_bsgs(hipStream_t s, FieldKind fk, uint32_t tbits, const PassBuffers& pb, uint32_t n_ent,
                 uint32_t grid) {
    const int W = tbits <= 12 ? 2 : 1;
    const size_t lds = (size_t)W * ((size_t)1 << tbits) * sizeof(uint32_t);
    if (n_ent) grid = (n_ent + W - 1) / W;
    if (grid == 0) return;
    if (W == 2) {
        if (fk == FIELD_F64)
            hipLaunchKernelGGL((k_bsgs<F64Field, 2>), dim3(grid), dim3(128), lds, s, pb, tbits);
        else
            hipLaunchKernelGGL((k_bsgs<MontField, 2>), dim3(grid), dim3(128), lds, s, pb, tbits);
    } else {
        if (fk == FIELD_F64)
            hipLaunchKernelGGL((k_bsgs<F64Field, 1>), dim3(grid), dim3(64), lds, s, pb, tbits);
        else
            hipLaunchKernelGGL((k_bsgs<MontField, 1>), dim3(grid), dim3(64), lds, s, pb, tbits);
    }
}

}  // namespace mps
