// Kernel launch interface of the MI355X sieve worker (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mps {

// Odd numbers per sieve segment: one workgroup marks one segment in an LDS bitmap.
constexpr uint32_t kSegOdds = 1u << 17;           // 16 KiB bitmap
constexpr uint32_t kSieveThreads = 1024;
constexpr uint32_t kCombPrimeBound = 64;          // sieving primes below this: register combs
constexpr uint32_t kCombPrimes = 17;              // odd primes below kCombPrimeBound (3 .. 61)
constexpr uint32_t kWavePrimeBound = 4096;        // below this (and from kCombPrimeBound): a wavefront per prime
constexpr uint32_t kEntWords = 72;                // u64 words per BSGS entry (576 B)
constexpr uint32_t kMaxClasses = 24;              // classes per prime in class mode
constexpr uint32_t kClassMaxM = 1024;             // default largest baby-step count in class mode
constexpr size_t kMaxBsgsLds = 64 * 1024;          // LDS of one BSGS workgroup at most
constexpr uint32_t kBsgsOversubscribe = 16;      // batch engine: BSGS grid / resident workgroups
constexpr uint32_t kCszEntries = 49 * 32;        // class-size table: (m, giant steps) by qp * 32 + s
constexpr uint32_t kSetupTile = 1024;             // candidates a class-mode setup workgroup sorts at a time
// Setup modes (ItemParams.qmode).
constexpr uint32_t kModePlain = 0, kModeClass = 1, kModeOrder = 2;
// Setup kernel variant of order mode with the medium primes of p - 1 (setup_variant); the variant
// without them is the order-mode kernel with none of their code.
constexpr uint32_t kModeOrderExt = 3;
// Order mode: every p - 1 is split by Qo = lcm(cm, B), B the longest of the prefix products
// 367567200 = 2^5 3^3 5^2 7 11 13 17, then times 19, 23, 29, 31, 37, with Qo below 2^52 (the CRT
// then stays exact in doubles).
constexpr uint64_t kOrderQ = 367567200ull;
constexpr uint32_t kOrderMorePrimes[5] = {19, 23, 29, 31, 37};
constexpr uint64_t kOrderQLimit = 1ull << 52;
constexpr uint32_t kOrderMaxPrimes = 16;          // distinct primes of Qo
// Order mode, medium primes: a prime r of p - 1 outside Qo (r from 19 up to the item's bound) whose
// power residue b^((p-1)/r) is not 1 fixes n modulo r as well (the discrete log of t^((p-1)/r) to
// that base, a search of at most r steps), dividing the class's exponent count by r.  The table
// holds the odd primes from 19 on with the constants of exact tests and remainders by them:
// r^-1 mod 2^64 and floor((2^64 - 1) / r) (x is a multiple of r exactly when x * r^-1 mod 2^64 is
// at most that bound), 2^32 mod r, 1/r rounded and r as a double (for x < 2^53, x is a multiple of
// r exactly when x - rint(x * (1/r)) * r, one FMA, is zero).
struct ExtPrime {
    uint32_t r, r32;
    double dinv;
    uint64_t rinv, lim;
    double rd;
};
constexpr uint32_t kExtPrimes = 256;              // 19 .. 1669; a table index fits 8 bits
constexpr uint32_t kExtTested = 3 * 64;           // primes an item may use: three per lane of a wavefront
constexpr uint32_t kExtDefaultBound = 1024;       // default largest medium prime
constexpr uint32_t kExtDefaultKmin = 131072;      // classes of fewer exponents are not narrowed
struct ExtTable {
    ExtPrime e[kExtPrimes];
};
constexpr ExtTable make_ext_table() {
    ExtTable t{};
    uint32_t n = 0;
    for (uint32_t r = 19; n < kExtPrimes; r += 2) {
        bool prime = true;
        for (uint32_t d = 3; d * d <= r; d += 2)
            if (r % d == 0) prime = false;
        if (!prime) continue;
        uint64_t inv = r;  // r * r == 1 (mod 8): 3 bits, doubled by each Newton step
        for (int k = 0; k < 5; ++k) inv *= 2 - (uint64_t)r * inv;
        t.e[n++] = ExtPrime{r, (uint32_t)((1ull << 32) % r), 1.0 / (double)r, inv, ~0ull / r, (double)r};
    }
    return t;
}
constexpr uint64_t kHostDecides = ~0ull;          // M of a record the host must decide exactly
// Relation records a pass of `odds` odd numbers can produce: every prime of the pass in the worst
// case (an item whose b^n + c vanishes for some in-class n has a relation at every prime), and
// primes are at most 0.29 of the odd numbers above 1000 (2 / ln 1000); 4096 more cover the primes
// below 1000, which the host decides.
constexpr uint64_t record_cap(uint64_t odds) { return odds / 3 + 4096; }

// Entry layout (u64 words).  B is the baby-step base, Bm = B^m the giant-step base.
enum EntryWord : uint32_t {
    E_P = 0,      // prime (or base-2 Fermat pseudoprime)
    E_AUX = 1,    // Montgomery -p^-1 mod 2^64, or the bits of 1/p (FP64 field)
    E_ONE = 2,    // Montgomery form of 1 (unused by the FP64 field)
    E_T = 3,      // target t = -c (field form): baby steps are t * B^i
    E_MG = 4,     // m (low 32) | giant steps per class (high 32)
    E_SSRC = 5,   // classes s (bits 0..7) | info/class mask (bits 8..31) | number index (bits 32..63)
    E_B64 = 6,    // B^64
    E_BM64 = 7,   // Bm^64
    E_BLO = 8,    // t * B^k, k = 0..7
    E_BHI = 16,   // B^(8k), k = 0..7
    E_MLO = 24,   // Bm^k, k = 0..7 (G * Bm^k with one class: E_META bit 12)
    E_MHI = 32,   // Bm^(8k), k = 0..7
    E_NB = 40,    // plain: n_start; class mode: N0
    E_META = 41,  // qp (bits 0..7) | stride 2 (bit 8) | odd k (bit 9) | class mode (bit 10) |
                  // order mode (bit 11: n = n_start + k * L, L in E_DJ) | G folded into E_MLO (bit 12)
    E_KLIM = 42,  // plain: jcount (solutions k = g*m - i must lie in [0, jcount))
    E_DJ = 43,    // class mode: dj of the kept classes, 8 bits each, 3 words (n_j = N0 + dj);
                  // order mode: the entry's exponent modulus L
    E_BL = 46,    // B^L, L = lanes of the BSGS workgroup (the baby-step stride of one lane)
    E_BML = 47,   // Bm^L (the giant-step stride of one lane)
    E_G = 48,     // giant start per class G_j = b^(n_j), j < s
};

// Stage counters.  Each counter that many workgroups update has a 128-byte line of its own, so
// atomics on different counters never contend for one L2 line.
struct Counters {
    uint32_t n_surv, p0[31];
    uint32_t n_cand, p1[31];
    uint32_t n_ent, p2[31];
    uint32_t n_out, n_flag2, n_fermat_fail, p3[29];
    unsigned long long w_fermat, w_setup, w_bsgs, p4[13];
};

// Per-item constants for the filter and setup stages.
// Jacobi symbol (x / p) of a fixed item value x = 2^e * a * sign, prepared on the host: with
// 0 < a < 2^16 the device needs p mod a (from r32 = 2^32 mod a) and a 16-bit Jacobi loop; a == 0
// (|x| of 2^16 or more, or x == 0) takes the general path.
struct JacobiPre {
    uint32_t a, r32;
    uint32_t neg;    // x < 0: the factor (-1/p)
    uint32_t e_odd;  // e odd: the factor (2/p)
    double dinv;     // 1.0 / a
};

struct ItemParams {
    int64_t b, c;
    JacobiPre jb, jt;       // for x = b and x = -c
    uint64_t N0, N1;        // N0 clamped at 0
    uint64_t cm, cres;      // class modulus (>= 1) and residue
    uint64_t n_start;       // first in-class n >= N0 (plain mode)
    uint64_t kcount;        // in-class n in [n_start, N1) (plain mode)
    uint64_t Qmax;          // class mode: lcm(cm, 24)
    uint32_t nsmodQ;        // class mode: n_start mod Qmax
    uint32_t qmode;         // setup mode: 0 plain, 1 per-prime classes, 2 per-prime order (kModeOrder)
    uint32_t cm_r32;        // 2^32 mod cm when cm < 2^16
    uint32_t Q_r32;         // 2^32 mod Qmax
    double cm_dinv;         // 1.0 / cm when cm < 2^16
    double Q_dinv;          // 1.0 / Qmax
    uint32_t mcap;          // plain mode: largest baby-step count (bounds the LDS table)
    uint32_t lanes;         // lanes of one BSGS workgroup
    // plain mode sizes by (stride - 1) * 2 + parity, valid when kcount <= p
    uint32_t pm[4], png[4];
    uint64_t pjc[4];
    // class mode sizes: csz[qp * 32 + s] = (m, giant steps per class)
    const uint2* csz;
    // order mode: Qo (above), 1 / Qo rounded, N0 mod Qo, the distinct prime factors of Qo and their
    // powers in Qo; per prime: the CRT basis element of Qo (1 modulo the prime's power, 0 modulo
    // the others), the prime's part of cm and cres modulo that part
    uint64_t Qo, n0modQo;
    double Qo_inv;
    uint32_t nopr;
    uint32_t opr[kOrderMaxPrimes], oprpow[kOrderMaxPrimes];
    uint64_t ocrt[kOrderMaxPrimes];
    uint32_t ocm[kOrderMaxPrimes], ocres[kOrderMaxPrimes];
    // per odd prime r of Qo: r^-1 mod 2^64 and floor((2^64 - 1) / r); x is a multiple of r exactly
    // when x * r^-1 mod 2^64 <= that bound, and the product is then x / r
    uint64_t oinv[kOrderMaxPrimes], olim[kOrderMaxPrimes];
    // order mode, medium primes: table indices [ext_first, ext_end) (ext_first: the first prime not
    // dividing Qo; empty: off) for classes of at least ext_kmin exponents, in the fields whose bit
    // (1 << FieldKind) is set in ext_fields
    uint32_t ext_first, ext_end;
    uint64_t ext_kmin;
    uint32_t ext_fields;
    // order mode in the Montgomery field, small bases (mont_pow2_small): window bits smw (0: off)
    // and the window powers |b|^(2^i), |c|^(2^i) (i < smw); b^e = (-1)^e |b|^e when bneg, and
    // t^e = (-c)^e = (-1)^e |c|^e when tneg (c > 0)
    uint32_t smw, bneg, tneg;
    uint32_t bsp[4], tsp[4];
};

// BSGS table geometry of a pass: 2^nbbits buckets of 8 slots; a slot holds a fingerprint of
// fpbits bits over a baby-step index of ibits bits.
struct TableGeom {
    uint32_t nbbits, fpbits, ibits, waves;
    uint32_t load;  // keys per bucket the table is sized for; an entry uses the buckets its m needs
};
// Everything a pass's kernels need besides the buffers, resident in device memory so that a
// captured graph of the pass chain can be replayed for any pass by rewriting this descriptor.
struct PassDesc {
    uint64_t lo;      // first odd number of the pass
    uint64_t nodd;    // odd numbers in the pass
    ItemParams ip;
    TableGeom tg;
};

struct PassBuffers {
    const uint64_t* vals;   // filter path: the numbers themselves (then surv is unused), else null
    uint32_t* surv;         // survivor odd indices relative to the pass start
    uint64_t* cand;         // candidate: index (low 32) | info (high 32)
    uint64_t* ent;          // entries, kEntWords u64 each
    uint64_t* out;          // records (number, M): M = smallest solution n, or kHostDecides
    Counters* cnt;
    uint32_t surv_cap, cand_cap, ent_cap, out_cap;
    uint32_t need_prime;    // 1: base-2 Fermat test first (sieve survivors may be composite)
    uint32_t out_index;     // 1: flagged records are indices into vals, else the numbers
};

enum FieldKind : int { FIELD_MONT = 0, FIELD_F64 = 1 };

// The setup kernel variant of a pass in field fk: kModeOrderExt for order mode with medium primes
// in that field, else the item's mode.
inline uint32_t setup_variant(const ItemParams& ip, FieldKind fk) {
    const bool ext = ip.qmode == kModeOrder && ip.ext_end > ip.ext_first && ((ip.ext_fields >> fk) & 1u);
    return ext ? kModeOrderExt : ip.qmode;
}

// One odd sieving prime with the constants of remainders by it: q, 2^32 mod q and 1/q (rounded).
struct SievePrime {
    uint32_t q, r32;
    double inv;
};
// The odd primes 3 .. 65521 in ascending order: n in all, the first n_comb below kCombPrimeBound, the
// first n_wave below kWavePrimeBound.
struct SieveTable {
    const SievePrime* sp;
    uint32_t n, n_comb, n_wave;
};

void launch_reset(hipStream_t s, Counters* cnt, uint32_t n_surv0);
void launch_sieve(hipStream_t s, const PassDesc* d, uint64_t nodd, const SieveTable& tab, const PassBuffers& pb,
                  uint32_t grid_override);
void launch_filter(hipStream_t s, const PassDesc* d, const PassBuffers& pb, uint32_t grid);
void launch_setup(hipStream_t s, FieldKind fk, uint32_t mode, const PassDesc* d, const PassBuffers& pb, uint32_t grid);
TableGeom table_geom(uint64_t m_max, uint64_t waves, uint32_t load);
constexpr uint32_t kBsgsWaves = 4;                // default wavefronts of one BSGS workgroup (one entry)
size_t bsgs_lds_bytes(const TableGeom& tg);

// Miller-Rabin over the pass's relation records (after launch_bsgs, same stream): composite
// numbers are replaced by 0.
constexpr uint32_t kPrimeRecordBlocks = 64;  // a pass may have a record per prime; 64 workgroups keep a short launch
void launch_prime_records(hipStream_t s, const PassBuffers& pb);
void launch_bsgs(hipStream_t s, FieldKind fk, const TableGeom& tg, const PassDesc* d, const PassBuffers& pb,
                 uint32_t n_ent, uint32_t grid);
// Workgroups of this BSGS variant resident on the device at once (cus compute units): the grid
// of a launch that drains however many entries the pass produced.
uint32_t bsgs_persistent_grid(FieldKind fk, const TableGeom& tg, int cus);
// k_bsgs addresses its table by LDS byte offset from 0: true when no variant holds static LDS (then
// the dynamic LDS, which holds the table, starts at address 0).
bool bsgs_lds_at_zero();

// One pass as a fixed chain: descriptor upload, reset, sieve, filter, setup, BSGS, counter and
// record download.  Every grid is fixed (the kernels read their element counts on the device).
struct PassChain {
    PassDesc* d_desc;
    const PassDesc* h_desc;   // pinned; read when the chain runs
    Counters* h_cnt;          // pinned
    uint64_t* h_out;          // pinned, out_words words
    size_t out_words;
    SieveTable tab;
    PassBuffers pb;
    uint32_t grid_sieve, grid_stage, grid_bsgs;
    FieldKind fk;
    uint32_t mode;  // setup mode of the item (kModePlain / kModeClass / kModeOrder)
    TableGeom tg;
};
// Enqueues the chain on stream s.
hipError_t enqueue_pass_chain(hipStream_t s, const PassChain& c);
// Builds the chain as a graph of explicit nodes (no stream capture).
hipError_t build_pass_graph(const PassChain& c, hipGraph_t* out);

}  // namespace mps
