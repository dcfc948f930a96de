// Host side of the MI355X GPU worker: the C-ABI of include/mps_gpu.h.
//
// One handle owns one HIP stream and one set of pass buffers.  sieve_filter cuts [lo, hi) into
// passes of up to pass_odds odd numbers; each pass runs reset -> sieve -> filter -> setup -> bsgs
// and reads the stage counts back to size the next launch (this is the plain per-chunk worker;
// the batched worker in mps_batch.hip keeps the counts on the device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "mps_gpu.h"
#include "mps_kernels.h"
#include "mps_worker_internal.h"

using namespace mps;

namespace {

uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

void write_str(char* dst, size_t cap, const std::string& s) {
    if (dst == nullptr || cap == 0) return;
    const size_t n = std::min(cap - 1, s.size());
    std::memcpy(dst, s.data(), n);
    dst[n] = 0;
}

uint64_t ceil_div_u(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Plain mode: the number of baby steps m (a multiple of the workgroup's lanes, at most mcap) and
// giant steps ng = ceil(jc / m) + 1 that need the fewest workgroup rounds.
void plain_size(uint64_t jc, uint32_t mcap, uint32_t L, uint32_t& m, uint32_t& ng) {
    uint64_t best_r = ~0ull, best_steps = ~0ull;
    for (uint64_t k = 1;; ++k) {
        const uint64_t mm = std::min<uint64_t>(k * L, mcap);
        const uint64_t g = ceil_div_u(jc, mm) + 1;
        const uint64_t r = ceil_div_u(mm, L) + ceil_div_u(g, L);
        if (r < best_r || (r == best_r && mm + g < best_steps)) {
            best_r = r;
            best_steps = mm + g;
            m = (uint32_t)mm;
            ng = (uint32_t)g;
        }
        if (mm >= mcap || mm >= jc) break;
    }
}

// Class mode: Gc = L * cr giant steps per class and m <= maxm baby steps with
// (Gc - 1) * m >= K - 1 (every k in [0, K) is g*m - i for some g < Gc, i < m), chosen for the
// fewest workgroup rounds ceil(m / L) + s * cr.
void class_size(uint64_t K, uint32_t s, uint32_t L, uint32_t maxm, uint32_t& m, uint32_t& gc) {
    uint64_t best = ~0ull, best_steps = ~0ull;
    for (uint64_t cr = 1;; ++cr) {
        if (best != ~0ull && s * cr > best) break;
        const uint64_t G = (uint64_t)L * cr;
        const uint64_t mm = K <= 1 ? 1 : std::max<uint64_t>(1, ceil_div_u(K - 1, G - 1));
        if (mm > maxm) continue;
        const uint64_t r = ceil_div_u(mm, L) + s * cr;
        if (r < best || (r == best && mm + s * G < best_steps)) {
            best = r;
            best_steps = mm + s * G;
            m = (uint32_t)mm;
            gc = (uint32_t)G;
        }
    }
}

// ------------------------------------------------------------------ allocation-free scratch
//
// mps_gpu_sieve_relations_noalloc runs one pass of at most kScratchOdds odd numbers on buffers in
// `scratch`, a file-scope device array (one copy per GPU, held until the process ends): stage
// counters, survivors, candidates, BSGS entries and records, laid out at 256-byte boundaries with
// the capacities the allocating path gives a pass of the same size.  Host code never uses the
// array's name as an address: each device's address comes from hipGetSymbolAddress at
// mps_gpu_create and is kept per device.
constexpr uint64_t kScratchOdds = MPS_NOALLOC_MAX_INTS / 2;
constexpr uint64_t kScratchEnt = kScratchOdds / 8;
constexpr uint64_t kScratchRec = mps::record_cap(kScratchOdds);
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kScrCnt = 0;
constexpr size_t kScrSurv = align256(kScrCnt + sizeof(mps::Counters));
constexpr size_t kScrCand = align256(kScrSurv + kScratchOdds * sizeof(uint32_t));
constexpr size_t kScrEnt = align256(kScrCand + kScratchOdds * sizeof(uint64_t));
constexpr size_t kScrRec = align256(kScrEnt + kScratchEnt * mps::kEntWords * sizeof(uint64_t));
constexpr size_t kScratchBytes = align256(kScrRec + kScratchRec * 2 * sizeof(uint64_t));
static_assert(kScratchBytes < (64ull << 20), "the scratch array stays modest");
constexpr size_t kStageRecords = 4096;  // pinned staging for the records, allocated at create
static_assert(kStageRecords <= kScratchRec, "the first slice is read with the counts");
constexpr int kMaxScratchDevices = 64;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};

// This device's address of `scratch` (the current device must be `device`), looked up once per
// device; nullptr if the lookup fails.
void* scratch_address(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> l(g_scratch_mu);
    if (g_scratch_dev[device] == nullptr) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_dev[device] = p;
    }
    return g_scratch_dev[device];
}

// The pass buffers of the scratch array at device address base.
mps::PassBuffers scratch_buffers(void* base) {
    uint8_t* b = static_cast<uint8_t*>(base);
    mps::PassBuffers pb{};
    pb.vals = nullptr;
    pb.cnt = reinterpret_cast<mps::Counters*>(b + kScrCnt);
    pb.surv = reinterpret_cast<uint32_t*>(b + kScrSurv);
    pb.cand = reinterpret_cast<uint64_t*>(b + kScrCand);
    pb.ent = reinterpret_cast<uint64_t*>(b + kScrEnt);
    pb.out = reinterpret_cast<uint64_t*>(b + kScrRec);
    pb.surv_cap = (uint32_t)kScratchOdds;
    pb.cand_cap = (uint32_t)kScratchOdds;
    pb.ent_cap = (uint32_t)kScratchEnt;
    pb.out_cap = (uint32_t)kScratchRec;
    pb.need_prime = 1;
    pb.out_index = 0;
    return pb;
}

}  // namespace

namespace mps {

// gpu_classes (Math.hpp:324-337): lcm(cm, 24) <= 48 and a long range; the reference also
// requires an even cm (class_policy 0).  The split is exact for any cm: Q is even and so is
// p - 1, so g = gcd(Q, p - 1) and qp = lcm(cm, g) are even and every class keeps one parity of n.
bool class_mode(const mps_item* it, bool classes_on, int class_policy, uint64_t& Q) {
    const uint64_t cm = it->cls_mod == 0 ? 1 : it->cls_mod;
    Q = cm / gcd_u64(cm, 24) * 24;
    if (!classes_on || it->N0 < 0 || it->N1 <= it->N0) return false;
    const uint64_t span = (uint64_t)(it->N1 - it->N0);
    return (class_policy == 1 || (cm & 1) == 0) && Q <= 48 && span / Q >= 4096;
}

// Distinct primes of Qo and their powers in Qo, smallest first.  False when the kernels cannot
// take Qo: a prime or a prime power of 2^16 or more, or more than kOrderMaxPrimes primes.
bool order_factors(uint64_t Qo, uint32_t* opr, uint32_t* oprpow, uint32_t& nopr) {
    nopr = 0;
    uint64_t q = Qo;
    for (uint32_t r = 2; q > 1 && r < 65536; ++r) {
        if (q % r) continue;
        uint32_t ra = 1;
        while (q % r == 0) {
            q /= r;
            if ((uint64_t)ra * r >= 65536) return false;
            ra *= r;
        }
        if (nopr == kOrderMaxPrimes) return false;
        opr[nopr] = r;
        oprpow[nopr++] = ra;
    }
    return q == 1;
}

// Order mode (class policy 2, classes on): one class per prime from the power residues of b and t
// modulo Qo = lcm(cm, B), B the longest prefix product of kOrderQ and kOrderMorePrimes that keeps
// Qo below 2^52.  A class modulus whose Qo the kernels cannot take (order_factors) leaves the item
// to class or plain mode.
bool order_mode(const mps_item* it, bool classes_on, int class_policy, uint64_t& Qo) {
    const uint64_t cm = it->cls_mod == 0 ? 1 : it->cls_mod;
    auto lcm_cm = [cm](uint64_t b) { return (unsigned __int128)(cm / gcd_u64(cm, b)) * b; };
    Qo = 0;
    uint64_t B = kOrderQ;
    if (lcm_cm(B) >= kOrderQLimit) return false;
    uint64_t q = (uint64_t)lcm_cm(B);
    for (uint32_t r : kOrderMorePrimes) {  // the longest prefix that keeps Qo below 2^52
        B *= r;
        const unsigned __int128 qn = lcm_cm(B);
        if (qn >= kOrderQLimit) break;
        q = (uint64_t)qn;
    }
    uint32_t opr[kOrderMaxPrimes], oprpow[kOrderMaxPrimes], nopr = 0;
    if (!order_factors(q, opr, oprpow, nopr)) return false;
    Qo = q;
    return classes_on && class_policy == 2 && it->N0 >= 0 && it->N1 > it->N0;
}

// The per-item constants of (x/p) for x = (neg ? -1 : 1) * mag.
JacobiPre jacobi_pre(bool neg, uint64_t mag) {
    JacobiPre J{0, 0, 0, 0};
    if (mag == 0) return J;
    uint64_t ax = mag;
    const uint32_t e = (uint32_t)__builtin_ctzll(ax);
    ax >>= e;
    if (ax >= 65536) return J;
    J.a = (uint32_t)ax;
    J.r32 = (uint32_t)((1ull << 32) % ax);
    J.dinv = 1.0 / (double)ax;
    J.neg = neg ? 1u : 0u;
    J.e_odd = e & 1u;
    return J;
}

bool item_params(const mps_item* it, bool classes_on, int class_policy, uint32_t mcap, uint32_t cmax,
                 uint32_t waves, uint32_t ext_bound, ItemParams& ip, uint64_t& m_max, uint2* csz, std::string& why) {
    std::memset(&ip, 0, sizeof(ip));
    ip.b = it->b;
    ip.c = it->c;
    const uint64_t N0 = it->N0 < 0 ? 0 : (uint64_t)it->N0;
    const uint64_t N1 = it->N1 < 0 ? 0 : (uint64_t)it->N1;
    const uint64_t cm = it->cls_mod == 0 ? 1 : it->cls_mod;
    const uint32_t L = 64 * waves;
    ip.N0 = N0;
    ip.N1 = N1;
    ip.cm = cm;
    ip.cres = it->cls_res;
    ip.n_start = N0 + ((it->cls_res % cm) + cm - (N0 % cm)) % cm;
    ip.kcount = ip.n_start >= N1 ? 0 : (N1 - ip.n_start + cm - 1) / cm;
    // b and t = -c by sign and magnitude (|INT64_MIN| = 2^63 is a magnitude like any other)
    ip.jb = jacobi_pre(it->b < 0, it->b < 0 ? 0ull - (uint64_t)it->b : (uint64_t)it->b);
    ip.jt = jacobi_pre(it->c > 0, it->c < 0 ? 0ull - (uint64_t)it->c : (uint64_t)it->c);
    ip.cm_r32 = cm < 65536 ? (uint32_t)((1ull << 32) % cm) : 0;
    ip.cm_dinv = cm < 65536 ? 1.0 / (double)cm : 0.0;
    ip.mcap = mcap;
    ip.lanes = L;
    uint64_t Q = 0, Qo = 0;
    const bool omode = order_mode(it, classes_on, class_policy, Qo);
    const bool qmode = !omode && class_mode(it, classes_on, class_policy == 2 ? 1 : class_policy, Q);
    if (qmode) {
        ip.Qmax = Q;
        ip.nsmodQ = (uint32_t)(ip.n_start % Q);
        ip.Q_r32 = (uint32_t)((1ull << 32) % Q);
        ip.Q_dinv = 1.0 / (double)Q;
    }
    ip.qmode = omode ? kModeOrder : qmode ? kModeClass : kModePlain;
    m_max = 1;
    for (uint32_t i = 0; i < kCszEntries; ++i) csz[i] = make_uint2(1, L);
    if (omode) {
        ip.Qo = Qo;
        ip.Qo_inv = 1.0 / (double)Qo;
        ip.n0modQo = N0 % Qo;
        if (!order_factors(Qo, ip.opr, ip.oprpow, ip.nopr)) {  // order_mode checked them
            why = "order mode: class modulus the kernels cannot take";
            return false;
        }
        for (uint32_t j = 0; j < ip.nopr; ++j) {
            const uint64_t qr = ip.oprpow[j], M = Qo / qr;
            uint64_t inv = 1;  // M^-1 mod qr by search (qr < 2^16)
            while ((M % qr) * inv % qr != 1 % qr) ++inv;
            ip.ocrt[j] = (uint64_t)((unsigned __int128)M * inv % Qo);
            uint64_t cr = 1;
            while (cm % (cr * ip.opr[j]) == 0) cr *= ip.opr[j];
            ip.ocm[j] = (uint32_t)cr;
            ip.ocres[j] = (uint32_t)(ip.cres % cr);
            uint64_t rinv = 1;
            if (ip.opr[j] & 1) {
                rinv = ip.opr[j];  // r * r == 1 (mod 8): 3 bits, doubled by each Newton step
                for (int k = 0; k < 5; ++k) rinv *= 2 - ip.opr[j] * rinv;
            }
            ip.oinv[j] = rinv;
            ip.olim[j] = ~0ull / ip.opr[j];
        }
        // medium primes: from the first table prime that does not divide Qo up to the bound, at most
        // kExtTested of them (a table index of at most 254, so that a pair slot never reads as a
        // free one); ext_bound 0 is the default bound in the Montgomery field only (in the FP64
        // field the narrowing costs the setup about what it saves the BSGS), 1 is none
        static constexpr ExtTable kExt = make_ext_table();
        const uint32_t bound = ext_bound == 0 ? kExtDefaultBound : ext_bound;
        uint32_t e0 = 0, e1 = 0;
        while (e0 < kExtPrimes - 1 && Qo % kExt.e[e0].r == 0) ++e0;
        for (e1 = e0; e1 < kExtPrimes - 1 && e1 < e0 + kExtTested && kExt.e[e1].r <= bound; ++e1) {
        }
        ip.ext_first = e0;
        ip.ext_end = e1;
        ip.ext_kmin = kExtDefaultKmin;
        ip.ext_fields = ext_bound == 0 ? 1u << FIELD_MONT : (1u << FIELD_MONT) | (1u << FIELD_F64);
        // small bases: the widest window (at most 4 bits) whose digit powers of |b| and |c| stay
        // below 2^24
        const uint64_t ab = it->b < 0 ? 0ull - (uint64_t)it->b : (uint64_t)it->b;
        const uint64_t ac = it->c < 0 ? 0ull - (uint64_t)it->c : (uint64_t)it->c;
        const uint64_t smax = std::max(ab, ac);
        ip.smw = 0;
        if (ab != 0 && ac != 0 && smax < (1ull << 24))
            for (uint32_t w = 4; w >= 1 && ip.smw == 0; --w) {
                uint64_t x = 1;
                bool fits = true;
                for (uint32_t k = 0; k + 1 < (1u << w) && fits; ++k) fits = (x *= smax) < (1ull << 24);
                if (fits) ip.smw = w;
            }
        ip.bneg = it->b < 0;
        ip.tneg = it->c > 0;
        for (uint32_t i = 0; i < 4; ++i) {
            ip.bsp[i] = 0;
            ip.tsp[i] = 0;
        }
        for (uint32_t i = 0; i < ip.smw; ++i) {
            uint64_t xb = ab, xc = ac;
            for (uint32_t k = 0; k < i; ++k) {
                xb *= xb;
                xc *= xc;
            }
            ip.bsp[i] = (uint32_t)xb;
            ip.tsp[i] = (uint32_t)xc;
        }
        // an entry's class has at most kcount members: m <= sqrt(kcount) in whole rounds, <= mcap
        uint64_t m = 1;
        while (m * m < ip.kcount) ++m;
        m = (m + L - 1) / L * L;
        m_max = std::max<uint64_t>(1, std::min<uint64_t>(m, mcap));
    } else if (qmode) {
        const uint64_t span = N1 - N0;
        for (uint32_t qp = 1; qp <= 48; ++qp) {
            if (ip.Qmax % qp != 0 || qp % cm != 0) continue;
            const uint64_t K = ceil_div_u(span, qp);
            for (uint32_t sc = 1; sc <= qp / cm; ++sc) {
                uint32_t m = 1, gc = L;
                class_size(K, sc, L, cmax, m, gc);
                csz[qp * 32 + sc] = make_uint2(m, gc);
                m_max = std::max<uint64_t>(m_max, m);
            }
        }
    } else if (ip.kcount > 0) {
        for (uint32_t combo = 0; combo < 4; ++combo) {
            const uint64_t stride = combo / 2 + 1, parity = combo & 1;
            if (ip.kcount <= parity) continue;
            const uint64_t jc = (ip.kcount - parity + stride - 1) / stride;
            uint32_t m = 1, ng = 2;
            plain_size(jc, mcap, L, m, ng);
            ip.pm[combo] = m;
            ip.png[combo] = ng;
            ip.pjc[combo] = jc;
            m_max = std::max<uint64_t>(m_max, m);
        }
        // primes below the exponent count cap it at p (run_bsgs, Math.hpp:398-400): m <= sqrt bound
        const uint64_t kmax = (N1 > N0 ? (N1 - N0) / cm : 0) + 2;
        m_max = std::max<uint64_t>(m_max, std::min<uint64_t>((uint64_t)std::sqrt((double)kmax) + 2, mcap));
    }
    if (m_max > 4096) {
        why = "BSGS table would exceed the LDS budget";
        return false;
    }
    return true;
}

uint32_t tune_key(const mps_gpu* g, const mps_item* item, uint64_t P0, bool classes_on) {
    const uint64_t cm = item->cls_mod == 0 ? 1 : item->cls_mod;
    const uint64_t span = item->N1 > item->N0 && item->N0 >= 0 ? (uint64_t)(item->N1 - item->N0) : 0;
    const uint32_t kind = cm == 1 ? 0u : (cm & 1) ? 2u : 1u;
    const uint32_t pbits = 64 - __builtin_clzll(P0 | 1), sbits = 64 - __builtin_clzll(span / cm | 1);
    return (classes_on ? 1u : 0u) | kind << 2 | (pbits / 6) << 4 | (sbits / 2) << 10 |
           (g->field_mode == 0 ? 1u << 16 : 0u);
}

// The candidate shape the autotuner is timing, seen only by the autotuning thread.
thread_local const mps_tuning* t_candidate = nullptr;

mps_tuning item_tuning(const mps_gpu* g, const mps_item* item, uint64_t P0, bool classes_on) {
    mps_tuning t{g->waves, g->mcap, kClassMaxM, 2, (uint32_t)g->class_policy, g->ext_bound};
    std::lock_guard<std::mutex> l(g->tune_mu);
    const mps_tuning* o = nullptr;
    if (t_candidate) {
        o = t_candidate;
    } else {
        const auto f = g->tuned.find(tune_key(g, item, P0, classes_on));
        if (f != g->tuned.end()) o = &f->second;
    }
    if (o) {
        if (o->waves) t.waves = o->waves;
        if (o->mcap) t.mcap = o->mcap;
        if (o->cmax) t.cmax = o->cmax;
        if (o->load) t.load = o->load;
        if (o->policy) t.policy = o->policy;
        if (o->ext) t.ext = o->ext;
    }
    return t;
}

bool plan_item(const mps_gpu* g, const mps_item* item, uint64_t P0, bool classes_on, Plan& pl, std::string& why) {
    const mps_tuning t = item_tuning(g, item, P0, classes_on);
    uint32_t waves = t.waves;
    const int policy = (int)t.policy;
    if (waves == 0) {
        // measured on MI355X: two wavefronts per entry for the class tables (m <= 1024) and short
        // plain ranges; four when plain-mode entries run up to 2048 baby steps
        if (!item_params(item, classes_on, policy, t.mcap, t.cmax, 2, t.ext, pl.ip, pl.m_max, pl.csz.data(), why)) return false;
        // order-mode entries are mostly short (their classes are narrow), whatever the item's cap
        waves = pl.m_max > 1024 && pl.ip.qmode != kModeOrder ? 4 : 2;
    }
    if (!item_params(item, classes_on, policy, t.mcap, t.cmax, waves, t.ext, pl.ip, pl.m_max, pl.csz.data(), why)) return false;
    pl.tg = table_geom(pl.m_max, waves, t.load);
    if (bsgs_lds_bytes(pl.tg) > kMaxBsgsLds) {
        why = "BSGS table would exceed the LDS budget";
        return false;
    }
    return true;
}

FieldKind pass_field(const mps_gpu* g, uint64_t hi) {
    if (g->field_mode == 0) return FIELD_MONT;
    return hi <= (1ull << 50) ? FIELD_F64 : FIELD_MONT;
}

}  // namespace mps

namespace {

void set_err(mps_gpu* g, const std::string& what) {
    std::lock_guard<std::mutex> l(g->st_mu);
    g->err = what;
}

bool hip_fail(mps_gpu* g, hipError_t e, const char* what) {
    set_err(g, std::string(what) + ": " + hipGetErrorString(e));
    g->broken = true;  // sticky, like sv_fail (Gpu.hpp:982-987)
    return false;
}

#define HIPCHK(expr, what)                                  \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(g, e_, what); \
    } while (0)

// The calling thread's pass lane, locked in lk: a free lane, else (all busy) the next in turn once
// it frees.
PassLane& take_lane(mps_gpu* g, std::unique_lock<std::mutex>& lk) {
    // g->lanes changes only under every lane's lock; read before a lane is held, it may be stale,
    // which is harmless: lane 0 always serves, and a retired lane's buffers are allocated again
    // by ensure_buffers if a call still takes it
    const int n = std::max(1, std::min(g->lanes.load(), kPassLanes));
    for (int i = 0; i < n; ++i) {
        std::unique_lock<std::mutex> t(g->lane[i].mu, std::try_to_lock);
        if (t.owns_lock()) {
            lk = std::move(t);
            return g->lane[i];
        }
    }
    PassLane& L = g->lane[g->next_lane.fetch_add(1) % (uint32_t)n];
    lk = std::unique_lock<std::mutex>(L.mu);
    return L;
}

// Device bytes of one lane's pass buffers for `odds` odd numbers per pass (ensure_buffers).
uint64_t lane_bytes(uint64_t odds) {
    const uint64_t ent_cap = std::max<uint64_t>(odds / 8, 4096);
    return odds * (sizeof(uint32_t) + sizeof(uint64_t)) + ent_cap * kEntWords * sizeof(uint64_t) +
           2 * record_cap(odds) * sizeof(uint64_t);
}

// Every lane, taken in index order: no pass runs while the holder changes settings or times shapes.
struct AllLanes {
    std::unique_lock<std::mutex> l[kPassLanes];
    explicit AllLanes(mps_gpu* g) {
        for (int i = 0; i < kPassLanes; ++i) l[i] = std::unique_lock<std::mutex>(g->lane[i].mu);
    }
};

void free_buffers(PassLane& L) {
    hipFree(L.pb.surv);
    hipFree(L.pb.cand);
    hipFree(L.pb.ent);
    hipFree(L.pb.out);
    hipHostFree(L.h_out);
    L.pb.surv = nullptr;
    L.pb.cand = nullptr;
    L.pb.ent = nullptr;
    L.pb.out = nullptr;
    L.h_out = nullptr;
    L.buf_odds = 0;
}

bool ensure_buffers(mps_gpu* g, PassLane& L, uint64_t odds) {
    if (L.buf_odds >= odds) return true;
    free_buffers(L);
    const uint64_t ent_cap = std::max<uint64_t>(odds / 8, 4096);
    const uint64_t out_cap = record_cap(odds);  // records of 2 words
    HIPCHK(hipMalloc(&L.pb.surv, odds * sizeof(uint32_t)), "surv alloc");
    HIPCHK(hipMalloc(&L.pb.cand, odds * sizeof(uint64_t)), "cand alloc");
    HIPCHK(hipMalloc(&L.pb.ent, ent_cap * kEntWords * sizeof(uint64_t)), "entry alloc");
    HIPCHK(hipMalloc(&L.pb.out, 2 * out_cap * sizeof(uint64_t)), "out alloc");
    HIPCHK(hipHostMalloc((void**)&L.h_out, 2 * out_cap * sizeof(uint64_t), hipHostMallocDefault), "pinned out");
    L.pb.surv_cap = (uint32_t)odds;
    L.pb.cand_cap = (uint32_t)odds;
    L.pb.ent_cap = (uint32_t)ent_cap;
    L.pb.out_cap = (uint32_t)out_cap;
    L.buf_odds = odds;
    return true;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    return ms;
}

bool read_counts(mps_gpu* g, PassLane& L, const PassBuffers& pb) {
    HIPCHK(hipMemcpyAsync(L.h_cnt, pb.cnt, sizeof(Counters), hipMemcpyDeviceToHost, L.stream), "count read");
    HIPCHK(hipStreamSynchronize(L.stream), "stream sync");
    return true;
}

// Records copied back with a full pass's counts, under the same wait (a pass has a few; more
// are fetched by read_records).
constexpr uint32_t kRecPrefix = 1024;

// Counts and the first `pre` records of the pass just enqueued into h_rec, one wait.
bool read_counts_and_records(mps_gpu* g, PassLane& L, const PassBuffers& pb, uint64_t* h_rec, size_t pre) {
    HIPCHK(hipMemcpyAsync(L.h_cnt, pb.cnt, sizeof(Counters), hipMemcpyDeviceToHost, L.stream), "count read");
    if (pre)
        HIPCHK(hipMemcpyAsync(h_rec, pb.out, 2 * pre * sizeof(uint64_t), hipMemcpyDeviceToHost, L.stream),
               "record read");
    HIPCHK(hipStreamSynchronize(L.stream), "stream sync");
    return true;
}

// The pass's n_out records in L.h_out: those past the prefix read_counts_and_records brought.
bool read_records(mps_gpu* g, PassLane& L, uint32_t n_out) {
    const size_t pre = std::min<size_t>(kRecPrefix, L.pb.out_cap);
    if (n_out <= pre) return true;
    HIPCHK(hipMemcpyAsync(L.h_out + 2 * pre, L.pb.out + 2 * pre, 2 * (n_out - pre) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, L.stream),
           "record read");
    HIPCHK(hipStreamSynchronize(L.stream), "stream sync");
    return true;
}

// Workgroups for n items of per_block each, capped at 8 per CU: the stage kernels loop over tiles,
// so a capped grid keeps their per-workgroup atomics few.
uint32_t blocks_for(const mps_gpu* g, uint64_t n, uint32_t per_block) {
    const uint64_t b = (n + per_block - 1) / per_block;
    const uint64_t cap = 8ull * (uint64_t)std::max(1, g->num_cu);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

// Stage times of the pass just read back on lane L, up to stage `last` (0 sieve .. 3 BSGS).
void add_stage_times(mps_gpu* g, PassLane& L, int last) {
    const float t[4] = {elapsed(L.ev[0], L.ev[1]), last >= 1 ? elapsed(L.ev[1], L.ev[2]) : 0.f,
                        last >= 2 ? elapsed(L.ev[2], L.ev[3]) : 0.f, last >= 3 ? elapsed(L.ev[3], L.ev[4]) : 0.f};
    std::lock_guard<std::mutex> l(g->st_mu);
    g->st.ms_sieve += t[0];
    g->st.ms_filter += t[1];
    g->st.ms_setup += t[2];
    g->st.ms_bsgs += t[3];
    if (last >= 3) g->st.n_bsgs_launches++;
}

// One pass through the stage chain on the buffers pb, on lane L's stream, descriptor and events.
// stop_stage: 0 after the sieve, 1 after the filter, 2 after setup, 3 = full.  Leaves the final
// counts in L.h_cnt and, after a full pass, the first rec_prefix records in h_rec.
bool run_pass(mps_gpu* g, PassLane& L, const PassBuffers& pb, uint64_t* h_rec, size_t rec_prefix, uint64_t plo,
              uint64_t pc, const ItemParams& ip, FieldKind fk, const TableGeom& tg, int stop_stage) {
    hipStream_t s = L.stream;
    L.h_desc->lo = plo;
    L.h_desc->nodd = pc;
    L.h_desc->ip = ip;
    L.h_desc->tg = tg;
    HIPCHK(hipMemcpyAsync(L.d_csz, L.h_csz, kCszEntries * sizeof(uint2), hipMemcpyHostToDevice, s), "size table");
    HIPCHK(hipMemcpyAsync(L.d_desc, L.h_desc, sizeof(PassDesc), hipMemcpyHostToDevice, s), "pass descriptor");
    const PassDesc* d = L.d_desc;
    launch_reset(s, pb.cnt, pb.vals == nullptr ? 0u : (uint32_t)pc);
    HIPCHK(hipEventRecord(L.ev[0], s), "event");
    if (pb.vals == nullptr) launch_sieve(s, d, pc, g->tab, pb, 0);
    HIPCHK(hipGetLastError(), "sieve launch");
    HIPCHK(hipEventRecord(L.ev[1], s), "event");
    if (stop_stage == 0) {
        if (!read_counts(g, L, pb)) return false;
        add_stage_times(g, L, 0);
        return true;
    }
    // The filter and setup kernels read their element counts on the device, so their grids are
    // sized for the largest count and nothing is read back until the BSGS grid needs the entry
    // count.
    launch_filter(s, d, pb, blocks_for(g, pc, 256));
    HIPCHK(hipGetLastError(), "filter launch");
    HIPCHK(hipEventRecord(L.ev[2], s), "event");
    if (stop_stage == 1) {
        if (!read_counts(g, L, pb)) return false;
        add_stage_times(g, L, 1);
        return true;
    }
    // class mode sorts tiles of kSetupTile candidates, a workgroup a tile at a time; order mode runs
    // rounds of 256
    const uint32_t setup_grid = blocks_for(g, pc, ip.qmode == kModeClass ? kSetupTile : 256);
    launch_setup(s, fk, setup_variant(ip, fk), d, pb, setup_grid);
    HIPCHK(hipGetLastError(), "setup launch");
    HIPCHK(hipEventRecord(L.ev[3], s), "event");
    if (stop_stage == 2) {
        if (!read_counts(g, L, pb)) return false;
        add_stage_times(g, L, 2);
        return true;
    }
    // BSGS over the entries setup wrote, on a grid of several times the resident workgroups that
    // reads the entry count on the device (no read-back before it)
    launch_bsgs(s, fk, tg, d, pb, 0, bsgs_persistent_grid(fk, tg, g->num_cu));
    HIPCHK(hipGetLastError(), "bsgs launch");
    HIPCHK(hipEventRecord(L.ev[4], s), "event");
    // the records' Miller-Rabin test reads the record count on the device, so it is enqueued now
    // and the counts and records come back under one wait
    launch_prime_records(s, pb);
    HIPCHK(hipGetLastError(), "record test launch");
    if (!read_counts_and_records(g, L, pb, h_rec, rec_prefix)) return false;
    add_stage_times(g, L, 3);
    return true;
}

void accumulate(mps_gpu* g, PassLane& L, const PassBuffers& pb, uint64_t pc) {
    const Counters& c = *L.h_cnt;
    std::lock_guard<std::mutex> l(g->st_mu);
    g->st.passes++;
    g->st.odds += pc;
    g->st.survivors += c.n_surv;
    g->st.candidates += c.n_cand;
    g->st.entries += c.n_ent;
    g->st.flagged += c.n_out;
    g->st.flag2 += c.n_flag2 + (c.n_ent > pb.ent_cap ? c.n_ent - pb.ent_cap : 0);
    g->st.fermat_fail += c.n_fermat_fail;
    g->st.w_fermat += c.w_fermat;
    g->st.w_setup += c.w_setup;
    g->st.w_bsgs += c.w_bsgs;
}

// Item constants plus the class-size table, staged in lane L's pinned memory for run_pass to upload.
bool prepare_item(mps_gpu* g, PassLane& L, const mps_item* item, uint64_t P0, bool classes_on, Plan& pl) {
    std::string why;
    if (!plan_item(g, item, P0, classes_on, pl, why)) {
        set_err(g, why);
        return false;
    }
    std::memcpy(L.h_csz, pl.csz.data(), pl.csz.size() * sizeof(uint2));
    pl.ip.csz = L.d_csz;
    return true;
}

void record_shape(mps_gpu* g, FieldKind fk, const TableGeom& tg) {
    std::lock_guard<std::mutex> l(g->st_mu);
    g->st.field = (int)fk;
    g->st.tbits = (int)tg.nbbits;
}


}  // namespace

extern "C" {

mps_gpu* mps_gpu_create(int device_index, char* banner, size_t banner_cap) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        write_str(banner, banner_cap, "GPU disabled: no HIP device");
        return nullptr;
    }
    if (device_index < 0 || device_index >= ndev) device_index = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_index) != hipSuccess) {
        write_str(banner, banner_cap, "GPU disabled: hipGetDeviceProperties failed");
        return nullptr;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        write_str(banner, banner_cap, std::string("GPU disabled: kernels are built for gfx950, device is ") +
                                          prop.gcnArchName);
        return nullptr;
    }
    mps_gpu* g = new mps_gpu();
    g->device = device_index;
    g->name = prop.name;
    g->num_cu = prop.multiProcessorCount;
    g->mem_bytes = prop.totalGlobalMem;
    g->lanes = 2 * lane_bytes(g->pass_odds) <= g->mem_bytes / 4 * 3 ? kPassLanes : 1;
    bool ok = hipSetDevice(device_index) == hipSuccess;
    for (PassLane& L : g->lane) {  // streams at the default priority
        ok = ok && hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) == hipSuccess &&
             hipMalloc(&L.pb.cnt, sizeof(Counters)) == hipSuccess &&
             hipMalloc(&L.d_csz, kCszEntries * sizeof(uint2)) == hipSuccess &&
             hipMalloc(&L.d_desc, sizeof(PassDesc)) == hipSuccess &&
             hipHostMalloc((void**)&L.h_desc, sizeof(PassDesc), hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc((void**)&L.h_csz, kCszEntries * sizeof(uint2), hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc((void**)&L.h_cnt, sizeof(Counters), hipHostMallocDefault) == hipSuccess;
        for (int i = 0; ok && i < 6; ++i) ok = hipEventCreate(&L.ev[i]) == hipSuccess;
    }
    // odd sieving primes 3 .. 65521 with 2^32 mod q and 1/q
    std::vector<SievePrime> sp;
    {
        std::vector<uint8_t> comp(65536, 0);
        for (uint32_t q = 3; q < 65536; q += 2) {
            if (comp[q]) continue;
            sp.push_back(SievePrime{q, (uint32_t)((1ull << 32) % q), 1.0 / (double)q});
            for (uint32_t m = q * q; m < 65536; m += 2 * q) comp[m] = 1;
        }
    }
    auto count_below = [&sp](uint32_t bound) {
        return (uint32_t)(std::find_if(sp.begin(), sp.end(), [bound](const SievePrime& x) { return x.q >= bound; }) -
                          sp.begin());
    };
    g->tab.n = (uint32_t)sp.size();
    g->tab.n_comb = count_below(kCombPrimeBound);
    g->tab.n_wave = count_below(kWavePrimeBound);
    ok = ok && g->tab.n_comb == kCombPrimes &&
         hipMalloc(&g->d_sp, sp.size() * sizeof(SievePrime)) == hipSuccess &&
         hipMemcpy(g->d_sp, sp.data(), sp.size() * sizeof(SievePrime), hipMemcpyHostToDevice) == hipSuccess &&
         hipHostMalloc((void**)&g->h_stage, kStageRecords * 2 * sizeof(uint64_t), hipHostMallocDefault) == hipSuccess;
    if (ok) g->scratch = scratch_address(device_index);  // with the device current (hipSetDevice above)
    g->tab.sp = g->d_sp;
    if (!ok) {
        write_str(banner, banner_cap, "GPU disabled: HIP setup failed");
        mps_gpu_destroy(g);
        return nullptr;
    }
    if (!bsgs_lds_at_zero()) {
        write_str(banner, banner_cap, "GPU disabled: the BSGS kernel's table is not at LDS address 0");
        mps_gpu_destroy(g);
        return nullptr;
    }
    g->ok = true;
    char buf[512];
    std::snprintf(buf, sizeof(buf), "GPU enabled (HIP): %s (%s), %d CUs, global mem %zu MB", prop.name,
                  prop.gcnArchName, prop.multiProcessorCount, (size_t)(prop.totalGlobalMem >> 20));
    write_str(banner, banner_cap, buf);
    return g;
}

int mps_gpu_ok(const mps_gpu* g) { return g != nullptr && g->ok ? 1 : 0; }

void mps_gpu_set_classes(mps_gpu* g, int on) {
    if (g == nullptr) return;
    AllLanes all(g);
    g->classes_on = on != 0;
}

int mps_gpu_set_class_policy(mps_gpu* g, int policy) {
    if (g == nullptr || policy < 0 || policy > 2) return MPS_ERROR;
    AllLanes all(g);
    g->class_policy = policy;
    return MPS_OK;
}

int mps_gpu_configure(mps_gpu* g, int field, uint64_t pass_odds, uint32_t mcap, uint32_t waves) {
    if (g == nullptr) return MPS_ERROR;
    // every argument is checked before any takes effect
    if (field < -1 || field > 1) return MPS_ERROR;
    if (waves && waves != 1 && waves != 2 && waves != 4) return MPS_ERROR;
    if (mcap && (mcap < 256 || mcap > 4096)) return MPS_ERROR;
    if (pass_odds && (pass_odds < 4096 || pass_odds > (1ull << 31))) return MPS_ERROR;
    AllLanes all(g);
    g->field_mode = field;
    if (waves) g->waves = waves;
    if (mcap) g->mcap = mcap;
    if (pass_odds) {
        g->pass_odds = pass_odds;
        // a second lane only while two passes' buffers take at most 3/4 of the device's memory
        g->lanes = 2 * lane_bytes(pass_odds) <= g->mem_bytes / 4 * 3 ? kPassLanes : 1;
        if (g->lanes.load() < kPassLanes) {
            hipSetDevice(g->device);
            for (int i = g->lanes.load(); i < kPassLanes; ++i) free_buffers(g->lane[i]);
        }
    }
    return MPS_OK;
}

}  // extern "C"

namespace {

// Device records (number, M) of [lo|1, hi): the plain per-chunk pipeline, pass by pass.
int sieve_records(mps_gpu* g, PassLane& L, uint64_t lo, uint64_t hi, const mps_item* item, std::vector<uint64_t>& rec) {
    rec.clear();
    if (!g->ok || g->broken) return MPS_ERROR;
    if (hi >= (1ull << 62) || lo <= 3) {
        set_err(g, "range outside (3, 2^62)");
        return MPS_DECLINED;
    }
    lo |= 1;
    if (lo >= hi) return MPS_OK;
    Plan pl;
    if (!prepare_item(g, L, item, lo, g->classes_on, pl)) return MPS_DECLINED;
    const ItemParams& ip = pl.ip;
    const FieldKind fk = pass_field(g, hi);
    const TableGeom tg = pl.tg;
    const uint64_t count = (hi - lo + 1) / 2;
    if (hipSetDevice(g->device) != hipSuccess) return MPS_ERROR;
    if (!ensure_buffers(g, L, std::min(count, g->pass_odds))) return MPS_ERROR;
    L.pb.vals = nullptr;
    L.pb.need_prime = 1;
    L.pb.out_index = 0;
    for (uint64_t pb0 = 0; pb0 < count; pb0 += g->pass_odds) {
        const uint64_t pc = std::min(g->pass_odds, count - pb0);
        const uint64_t plo = lo + 2 * pb0;
        if (!run_pass(g, L, L.pb, L.h_out, std::min<size_t>(kRecPrefix, L.pb.out_cap), plo, pc, ip, fk, tg, 3))
            return MPS_ERROR;
        accumulate(g, L, L.pb, pc);
        const uint32_t n_out = L.h_cnt->n_out;
        if (n_out > L.pb.out_cap) {
            set_err(g, "device records exceed the output buffer");
            return MPS_DECLINED;
        }
        if (n_out) {
            if (!read_records(g, L, n_out)) return MPS_ERROR;
            for (uint32_t i = 0; i < n_out; ++i) {
                if (L.h_out[2 * i] == 0) continue;  // composite (k_prime_records)
                rec.push_back(L.h_out[2 * i]);
                rec.push_back(L.h_out[2 * i + 1]);
            }
        }
    }
    record_shape(g, fk, tg);
    return MPS_OK;
}

}  // namespace

extern "C" {

int mps_gpu_sieve_filter(mps_gpu* g, uint64_t lo, uint64_t hi, const mps_item* item, uint64_t* out,
                         size_t out_cap, size_t* out_n) {
    if (out_n) *out_n = 0;
    if (g == nullptr || item == nullptr || out_n == nullptr) return MPS_ERROR;
    std::unique_lock<std::mutex> lock;
    PassLane& L = take_lane(g, lock);
    std::vector<uint64_t> rec;
    const int st = sieve_records(g, L, lo, hi, item, rec);
    if (st != MPS_OK) return st;
    const size_t n = rec.size() / 2;
    for (size_t i = 0; i < n && i < out_cap && out; ++i) out[i] = rec[2 * i];
    *out_n = n;
    return n > out_cap ? MPS_CAPACITY : MPS_OK;
}

int mps_gpu_sieve_relations(mps_gpu* g, uint64_t lo, uint64_t hi, const mps_item* item, mps_dev_relation* out,
                            size_t out_cap, size_t* out_n) {
    if (out_n) *out_n = 0;
    if (g == nullptr || item == nullptr || out_n == nullptr) return MPS_ERROR;
    std::unique_lock<std::mutex> lock;
    PassLane& L = take_lane(g, lock);
    std::vector<uint64_t> rec;
    const int st = sieve_records(g, L, lo, hi, item, rec);
    if (st != MPS_OK) return st;
    const size_t n = rec.size() / 2;
    for (size_t i = 0; i < n && i < out_cap && out; ++i) {
        out[i].p = rec[2 * i];
        out[i].M = rec[2 * i + 1];
    }
    *out_n = n;
    return n > out_cap ? MPS_CAPACITY : MPS_OK;
}

int mps_gpu_sieve_relations_noalloc(mps_gpu* g, uint64_t lo, uint64_t hi, const mps_item* item,
                                    mps_dev_relation* out, size_t out_cap, size_t* out_n) {
    if (out_n) *out_n = 0;
    if (g == nullptr || item == nullptr || out_n == nullptr) return MPS_ERROR;
    std::lock_guard<std::mutex> one(g->noalloc_mu);  // the scratch array serves one call at a time
    std::unique_lock<std::mutex> lock;
    PassLane& L = take_lane(g, lock);  // its stream, descriptor and staging; the buffers are scratch
    if (!g->ok || g->broken) return MPS_ERROR;
    if (g->scratch == nullptr) {
        set_err(g, "no-allocation path: the scratch array's device address is unknown");
        return MPS_ERROR;
    }
    if (hi > lo && hi - lo > MPS_NOALLOC_MAX_INTS) {
        set_err(g, "no-allocation path: a range of more than 2^20 integers (MPS_NOALLOC_MAX_INTS)");
        return MPS_ERROR;
    }
    if (hi >= (1ull << 62) || lo <= 3) {
        set_err(g, "range outside (3, 2^62)");
        return MPS_DECLINED;
    }
    lo |= 1;
    if (lo >= hi) return MPS_OK;
    Plan pl;
    if (!prepare_item(g, L, item, lo, g->classes_on, pl)) return MPS_DECLINED;
    const FieldKind fk = pass_field(g, hi);
    const uint64_t count = (hi - lo + 1) / 2;  // at most kScratchOdds
    const PassBuffers pb = scratch_buffers(g->scratch);
    if (hipSetDevice(g->device) != hipSuccess) return MPS_ERROR;
    if (!run_pass(g, L, pb, g->h_stage, kStageRecords, lo, count, pl.ip, fk, pl.tg, 3)) return MPS_ERROR;
    accumulate(g, L, pb, count);
    const uint32_t n_rec = L.h_cnt->n_out;
    if (n_rec > pb.out_cap) {
        set_err(g, "device records exceed the scratch record area");
        return MPS_DECLINED;
    }
    // the records through the pinned staging area, kStageRecords at a time; composites (number 0,
    // k_prime_records) are dropped
    size_t n = 0;
    for (uint32_t off = 0; off < n_rec; off += (uint32_t)kStageRecords) {
        const uint32_t k = std::min<uint32_t>((uint32_t)kStageRecords, n_rec - off);
        if (off > 0) {  // the first slice came back with the counts
            hipError_t e = hipMemcpyAsync(g->h_stage, pb.out + 2 * (uint64_t)off, 2 * k * sizeof(uint64_t),
                                          hipMemcpyDeviceToHost, L.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
            if (e != hipSuccess) {
                hip_fail(g, e, "record read");
                return MPS_ERROR;
            }
        }
        for (uint32_t i = 0; i < k; ++i) {
            if (g->h_stage[2 * i] == 0) continue;
            if (n < out_cap && out) {
                out[n].p = g->h_stage[2 * i];
                out[n].M = g->h_stage[2 * i + 1];
            }
            ++n;
        }
    }
    record_shape(g, fk, pl.tg);
    *out_n = n;
    return n > out_cap ? MPS_CAPACITY : MPS_OK;
}

int mps_gpu_filter(mps_gpu* g, const uint64_t* primes, size_t n, const mps_item* item, uint64_t* todo,
                   size_t todo_cap, size_t* n_todo) {
    if (n_todo) *n_todo = 0;
    if (g == nullptr || item == nullptr || n_todo == nullptr) return MPS_ERROR;
    std::unique_lock<std::mutex> lock;
    PassLane& L = take_lane(g, lock);
    if (!g->ok || g->broken) return MPS_ERROR;
    if (n == 0) return MPS_OK;
    uint64_t pmax = 0, pmin = ~0ull;
    for (size_t i = 0; i < n; ++i) {
        pmax = std::max(pmax, primes[i]);
        pmin = std::min(pmin, primes[i]);
    }
    Plan pl;
    if (!prepare_item(g, L, item, pmin, false, pl)) return MPS_DECLINED;  // filter(): plain prep path
    const ItemParams& ip = pl.ip;
    if (pmax >= (1ull << 62)) {
        set_err(g, "prime above 2^62");
        return MPS_DECLINED;
    }
    const FieldKind fk = pass_field(g, pmax + 1);
    const TableGeom tg = pl.tg;
    if (hipSetDevice(g->device) != hipSuccess) return MPS_ERROR;
    if (!ensure_buffers(g, L, g->pass_odds)) return MPS_ERROR;
    // every value is a prime and may take an entry, so a batch is at most the entry list: no entry
    // overflows to the host, and the flagged set does not depend on the order of the device's
    // slot reservations
    const size_t batch = (size_t)std::min<uint64_t>(L.pb.ent_cap, 1ull << 22);
    if (L.vals_cap < batch) {
        hipFree(L.d_vals);
        L.d_vals = nullptr;
        if (hipMalloc(&L.d_vals, batch * sizeof(uint64_t)) != hipSuccess) {
            hip_fail(g, hipErrorOutOfMemory, "values alloc");
            return MPS_ERROR;
        }
        L.vals_cap = batch;
    }
    std::vector<uint64_t> idx;
    for (size_t off = 0; off < n; off += batch) {
        const size_t cnt = std::min(batch, n - off);
        if (hipMemcpyAsync(L.d_vals, primes + off, cnt * sizeof(uint64_t), hipMemcpyHostToDevice, L.stream) != hipSuccess) {
            hip_fail(g, hipErrorUnknown, "values upload");
            return MPS_ERROR;
        }
        L.pb.vals = L.d_vals;
        L.pb.need_prime = 0;
        L.pb.out_index = 1;
        const bool ok = run_pass(g, L, L.pb, L.h_out, std::min<size_t>(kRecPrefix, L.pb.out_cap), 0, cnt, ip, fk, tg, 3);
        L.pb.vals = nullptr;
        L.pb.need_prime = 1;
        L.pb.out_index = 0;
        if (!ok) return MPS_ERROR;
        accumulate(g, L, L.pb, 0);
        const uint32_t n_out = L.h_cnt->n_out;
        if (n_out > L.pb.out_cap) {
            set_err(g, "device records exceed the output buffer");
            return MPS_DECLINED;
        }
        if (n_out) {
            if (!read_records(g, L, n_out)) return MPS_ERROR;
            for (uint32_t i = 0; i < n_out; ++i) idx.push_back(off + L.h_out[2 * i]);
        }
    }
    std::sort(idx.begin(), idx.end());
    idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
    for (size_t i = 0; i < idx.size() && i < todo_cap; ++i) todo[i] = idx[i];
    *n_todo = idx.size();
    return idx.size() > todo_cap ? MPS_CAPACITY : MPS_OK;
}

int mps_gpu_debug_stage(mps_gpu* g, int stage, uint64_t lo, uint64_t hi, const mps_item* item, uint64_t* out,
                        size_t out_cap_words, size_t* out_n) {
    if (out_n) *out_n = 0;
    if (g == nullptr || item == nullptr || out_n == nullptr || stage < 0 || stage > 2) return MPS_ERROR;
    std::unique_lock<std::mutex> lock;
    PassLane& L = take_lane(g, lock);
    if (!g->ok || g->broken) return MPS_ERROR;
    lo |= 1;
    if (lo <= 3 || hi >= (1ull << 62) || lo >= hi) return MPS_DECLINED;
    Plan pl;
    if (!prepare_item(g, L, item, lo, g->classes_on, pl)) return MPS_DECLINED;
    const ItemParams& ip = pl.ip;
    const FieldKind fk = pass_field(g, hi);
    const TableGeom tg = pl.tg;
    const uint64_t count = (hi - lo + 1) / 2;
    if (count > g->pass_odds) {
        set_err(g, "debug stage covers one pass only");
        return MPS_DECLINED;
    }
    if (hipSetDevice(g->device) != hipSuccess || !ensure_buffers(g, L, count)) return MPS_ERROR;
    L.pb.vals = nullptr;
    L.pb.need_prime = 1;
    L.pb.out_index = 0;
    if (!run_pass(g, L, L.pb, L.h_out, std::min<size_t>(kRecPrefix, L.pb.out_cap), lo, count, ip, fk, tg, stage))
        return MPS_ERROR;
    const Counters c = *L.h_cnt;
    size_t words = 0, recs = 0;
    if (stage == 0) {
        const uint32_t n = std::min(c.n_surv, L.pb.surv_cap);
        std::vector<uint32_t> v(n);
        if (n && hipMemcpy(v.data(), L.pb.surv, n * 4, hipMemcpyDeviceToHost) != hipSuccess) return MPS_ERROR;
        for (uint32_t i = 0; i < n; ++i) {
            if (words < out_cap_words) out[words] = lo + 2 * (uint64_t)v[i];
            ++words;
        }
        recs = n;
    } else if (stage == 1) {
        const uint32_t n = std::min(c.n_cand, L.pb.cand_cap);
        std::vector<uint64_t> v(n);
        if (n && hipMemcpy(v.data(), L.pb.cand, n * 8, hipMemcpyDeviceToHost) != hipSuccess) return MPS_ERROR;
        for (uint32_t i = 0; i < n; ++i) {
            if (words + 1 < out_cap_words) {
                out[words] = lo + 2 * (v[i] & 0xFFFFFFFFull);
                out[words + 1] = v[i] >> 32;
            }
            words += 2;
        }
        recs = n;
    } else {
        const uint32_t n = std::min(c.n_ent, L.pb.ent_cap);
        std::vector<uint64_t> v((size_t)n * kEntWords);
        if (n && hipMemcpy(v.data(), L.pb.ent, v.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) return MPS_ERROR;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t* E = &v[(size_t)i * kEntWords];
            if (words + 3 < out_cap_words) {
                out[words] = E[E_P];
                out[words + 1] = E[E_MG] & 0xFFFFFFFFull;
                out[words + 2] = E[E_MG] >> 32;
                out[words + 3] = E[E_SSRC] & 0xFFFFFFFFull;
            }
            words += 4;
        }
        recs = n;
    }
    *out_n = recs;
    return words > out_cap_words ? MPS_CAPACITY : MPS_OK;
}

const char* mps_gpu_last_error(const mps_gpu* g) { return g ? g->err.c_str() : "null handle"; }

int mps_gpu_get_stats(const mps_gpu* g, mps_gpu_stats* s) {
    if (g == nullptr || s == nullptr) return MPS_ERROR;
    std::lock_guard<std::mutex> l(g->st_mu);
    *s = g->st;
    return MPS_OK;
}

void mps_gpu_reset_stats(mps_gpu* g) {
    if (g == nullptr) return;
    std::lock_guard<std::mutex> l(g->st_mu);
    g->st = mps_gpu_stats{};
}

void mps_gpu_destroy(mps_gpu* g) {
    if (g == nullptr) return;
    if (g->ok) hipSetDevice(g->device);
    for (PassLane& L : g->lane) {
        if (L.stream) hipStreamSynchronize(L.stream);
        free_buffers(L);
        hipFree(L.d_vals);
        hipFree(L.pb.cnt);
        hipHostFree(L.h_cnt);
        hipFree(L.d_csz);
        hipHostFree(L.h_csz);
        hipFree(L.d_desc);
        hipHostFree(L.h_desc);
        for (int i = 0; i < 6; ++i)
            if (L.ev[i]) hipEventDestroy(L.ev[i]);
        if (L.stream) hipStreamDestroy(L.stream);
    }
    hipFree(g->d_sp);
    hipHostFree(g->h_stage);
    delete g;
}

}  // extern "C"

namespace {

bool valid_tuning(const mps_tuning& t) {
    return (t.waves == 0 || t.waves == 1 || t.waves == 2 || t.waves == 4) && (t.mcap == 0 || (t.mcap >= 256 && t.mcap <= 4096)) &&
           (t.cmax == 0 || (t.cmax >= 256 && t.cmax <= 2048)) && (t.load == 0 || t.load == 2 || t.load == 4) &&
           t.policy <= 2;
}

double device_ms(const mps_gpu* g) {
    std::lock_guard<std::mutex> l(g->st_mu);
    return g->st.ms_sieve + g->st.ms_filter + g->st.ms_setup + g->st.ms_bsgs;
}

// One timed run of the sample under the current forced tuning: device ms, records sorted.
int timed_sample(mps_gpu* g, PassLane& L, uint64_t lo, uint64_t hi, const mps_item* item, std::vector<uint64_t>& rec,
                 double& ms) {
    const double before = device_ms(g);
    const int st = sieve_records(g, L, lo, hi, item, rec);
    ms = device_ms(g) - before;
    std::vector<std::pair<uint64_t, uint64_t>> pr(rec.size() / 2);
    for (size_t i = 0; i < pr.size(); ++i) pr[i] = {rec[2 * i], rec[2 * i + 1]};
    std::sort(pr.begin(), pr.end());
    for (size_t i = 0; i < pr.size(); ++i) {
        rec[2 * i] = pr[i].first;
        rec[2 * i + 1] = pr[i].second;
    }
    return st;
}

// Records of two runs agree: a number both decided has the same smallest solution, and a number
// one run decided to have a solution appears in the other (decided alike, or left to the host).
// A number may be left to the host by one shape and decided by another (the cycle test depends on
// the baby-step count), so neither side's host-decided numbers need a counterpart.
bool records_agree(const std::vector<uint64_t>& a, const std::vector<uint64_t>& b) {
    std::map<uint64_t, uint64_t> ma, mb;
    for (size_t i = 0; i + 1 < a.size(); i += 2) ma[a[i]] = a[i + 1];
    for (size_t i = 0; i + 1 < b.size(); i += 2) mb[b[i]] = b[i + 1];
    for (int pass = 0; pass < 2; ++pass) {
        const auto& x = pass ? mb : ma;
        const auto& y = pass ? ma : mb;
        for (const auto& kv : x) {
            if (kv.second == kHostDecides) continue;
            const auto f = y.find(kv.first);
            if (f == y.end() || (f->second != kHostDecides && f->second != kv.second)) return false;
        }
    }
    return true;
}

}  // namespace

extern "C" {

int mps_gpu_autotune(mps_gpu* g, const mps_item* item, uint64_t P0, uint64_t sample_odds, mps_tuning* best,
                     double* best_ms, double* default_ms) {
    if (g == nullptr || item == nullptr || sample_odds == 0) return MPS_ERROR;
    AllLanes all(g);  // no other pass shares the device while shapes are timed
    PassLane& L = g->lane[0];
    if (!g->ok || g->broken) return MPS_ERROR;
    const uint64_t lo = P0 | 1;
    sample_odds = std::min(sample_odds, g->pass_odds);
    const uint64_t hi = lo + 2 * sample_odds;
    if (lo <= 3 || hi >= (1ull << 62)) return MPS_DECLINED;
    mps_gpu_stats keep;
    {
        std::lock_guard<std::mutex> l(g->st_mu);
        keep = g->st;
    }
    auto restore_stats = [&]() {
        std::lock_guard<std::mutex> l(g->st_mu);
        g->st = keep;
    };
    mps_tuning cand{};
    auto force = [&](bool on, const mps_tuning& t) {
        cand = t;
        t_candidate = on ? &cand : nullptr;
    };
    // the defaults (no table entry for the key) give the reference records and time
    const uint32_t key = tune_key(g, item, P0, g->classes_on);
    force(true, mps_tuning{g->waves, g->mcap, kClassMaxM, 2, (uint32_t)g->class_policy, g->ext_bound});
    std::vector<uint64_t> ref, rec;
    double ms = 0, t_default = 1e300;
    int st = MPS_OK;
    for (int rep = 0; rep < 3 && st == MPS_OK; ++rep) {  // the first run warms the kernels up
        st = timed_sample(g, L, lo, hi, item, ref, ms);
        if (rep > 0) t_default = std::min(t_default, ms);
    }
    if (st != MPS_OK) {
        force(false, mps_tuning{});
        restore_stats();
        return st;
    }
    mps_tuning win{0, 0, 0, 0, 0, 0};
    double t_win = t_default;
    // times one candidate shape (skipped when its table would not fit); false on an error
    auto try_shape = [&](const mps_tuning& t) {
        Plan pl;
        std::string why;
        force(true, t);
        if (!plan_item(g, item, P0, g->classes_on, pl, why)) return true;  // the table would not fit
        double tbest = 1e300;
        for (int rep = 0; rep < 2 && st == MPS_OK; ++rep) {
            st = timed_sample(g, L, lo, hi, item, rec, ms);
            tbest = std::min(tbest, ms);
            if (st == MPS_OK && !records_agree(rec, ref)) {
                set_err(g, "autotune: a candidate shape contradicts the default's records");
                st = MPS_ERROR;
            }
        }
        if (st != MPS_OK) return false;
        if (tbest < t_win) {
            t_win = tbest;
            win = t;
        }
        return true;
    };
    // with classes on, both class policies that change the setup are candidates; policy 2 only
    // where the item has an order-mode reduction (otherwise it runs as policy 1)
    std::vector<uint32_t> policies{(uint32_t)g->class_policy};
    uint64_t Q = 0, Qo = 0;
    if (g->classes_on) {
        policies = {1u};
        if (order_mode(item, true, 2, Qo)) policies.push_back(2u);
    }
    for (uint32_t policy : policies) {
    const bool qmode = !order_mode(item, g->classes_on, (int)policy, Qo) &&
                       class_mode(item, g->classes_on, policy == 2 ? 1 : (int)policy, Q);
    for (uint32_t waves : {1u, 2u, 4u}) {
        for (uint32_t load : {2u, 4u}) {
            for (uint32_t cap : {0u, 1u, 2u, 3u}) {
                const mps_tuning t{waves, qmode ? g->mcap : 512u << cap, qmode ? 256u << cap : kClassMaxM, load, policy,
                                   g->ext_bound};
                if (!try_shape(t)) break;
            }
            if (st != MPS_OK) break;
        }
        if (st != MPS_OK) break;
    }
    if (st != MPS_OK) break;
    }
    // the medium primes of p - 1 off and on (in both fields) for the winning shape, in order mode
    if (st == MPS_OK) {
        mps_tuning base = win;
        if (base.waves == 0 && base.mcap == 0 && base.load == 0 && base.policy == 0)  // the defaults won
            base = mps_tuning{g->waves, g->mcap, kClassMaxM, 2, (uint32_t)g->class_policy, g->ext_bound};
        if (order_mode(item, g->classes_on, (int)base.policy, Qo))
            for (uint32_t ext : {1u, kExtDefaultBound}) {
                mps_tuning t = base;
                t.ext = ext;
                if (!try_shape(t)) break;
            }
    }
    force(false, mps_tuning{});
    restore_stats();
    if (st != MPS_OK) return st;
    {
        std::lock_guard<std::mutex> l(g->tune_mu);
        g->tuned[key] = win;  // all zero: the defaults were fastest
    }
    if (best) *best = item_tuning(g, item, P0, g->classes_on);
    if (best_ms) *best_ms = t_win;
    if (default_ms) *default_ms = t_default;
    return MPS_OK;
}

int mps_gpu_get_tuning(mps_gpu* g, const mps_item* item, uint64_t P0, mps_tuning* t, uint32_t* key) {
    if (g == nullptr || item == nullptr || t == nullptr) return MPS_ERROR;
    *t = item_tuning(g, item, P0, g->classes_on);
    if (key) *key = tune_key(g, item, P0, g->classes_on);
    return MPS_OK;
}

int mps_gpu_set_tuning(mps_gpu* g, const mps_item* item, uint64_t P0, const mps_tuning* t) {
    if (g == nullptr || item == nullptr || t == nullptr || !valid_tuning(*t)) return MPS_ERROR;
    std::lock_guard<std::mutex> l(g->tune_mu);
    g->tuned[tune_key(g, item, P0, g->classes_on)] = *t;
    return MPS_OK;
}

void mps_gpu_clear_tuning(mps_gpu* g) {
    if (g == nullptr) return;
    std::lock_guard<std::mutex> l(g->tune_mu);
    g->tuned.clear();
}

}  // extern "C"
