// CPU check of the device arithmetic (mass-prime-siever_amd/csrc/mps_device.hpp), compiled for the
// host from the same header the kernels use.  Every result is compared with arithmetic that shares
// no code with it: unsigned __int128 products and remainders, exact integer properties of the
// doubles, and Euler's criterion for Jacobi symbols.  Prints one line per family and exits 1 on
// the first mismatch.  Usage: device_arith_check [scale] (scale 1 = the CPU-suite size).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mps_device.hpp"
#include "mps_kernels.h"

using namespace mps;
typedef unsigned __int128 u128;

static uint64_t g_rng = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {  // splitmix64
    uint64_t z = (g_rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t rnd_below(uint64_t n) { return n ? (uint64_t)((u128)rnd() * n >> 64) : 0; }

static long g_checks = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        ++g_checks;                                        \
        if (!(cond)) {                                     \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                      \
            std::printf("\n");                             \
            std::exit(1);                                  \
        }                                                  \
    } while (0)

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((u128)a * b % p); }
static uint64_t powmod(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) r = mulmod(r, a, p);
        a = mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}
// 12-base strong probable-prime test with 128-bit products: exact below 2^64.
static bool is_prime(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t B[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t b : B)
        if (n % b == 0) return n == b;
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) d >>= 1, ++s;
    for (uint64_t b : B) {
        uint64_t x = powmod(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool ok = false;
        for (int r = 1; r < s && !ok; ++r) ok = (x = mulmod(x, x, n)) == n - 1;
        if (!ok) return false;
    }
    return true;
}
static uint64_t random_prime(uint64_t lo, uint64_t hi) {
    for (;;) {
        uint64_t x = (lo + rnd_below(hi - lo)) | 1;
        if (x < hi && is_prime(x)) return x;
    }
}

// ------------------------------------------------------------------ Montgomery field, p < 2^62
static void check_mont(int scale) {
    std::vector<uint64_t> mods = {3, 5, 7, 65537, 4294967291ull, 4294967311ull, (1ull << 50) - 27, (1ull << 50) + 55,
                                  (1ull << 61) - 1, (1ull << 62) - 57, (1ull << 62) - 1, (1ull << 62) - 3,
                                  4000000000000000037ull};
    for (int i = 0; i < 200 * scale; ++i) mods.push_back((rnd_below(1ull << (2 + rnd_below(61))) | 1) | 3);
    long products = 0;
    for (uint64_t p : mods) {
        if (p >= (1ull << 62) || p < 3) continue;
        MontField f;
        f.init(p);
        const uint64_t R = (uint64_t)(((u128)1 << 64) % p);
        CHECK(f.one() == R, "one p=%llu", (unsigned long long)p);
        CHECK(f.r2() == (uint64_t)((u128)R * R % p), "r2 p=%llu", (unsigned long long)p);
        std::vector<uint64_t> ops = {0, 1, p - 1, p / 2, R};
        for (int k = 0; k < 64; ++k) ops.push_back(rnd_below(p));
        for (uint64_t a : ops)
            for (uint64_t b : ops) {
                const uint64_t x = f.mul(a, b);  // a * b * 2^-64 mod p: x * 2^64 == a * b (mod p)
                CHECK(x < p && (uint64_t)(((u128)x << 64) % p) == mulmod(a, b, p), "mont mul p=%llu a=%llu b=%llu",
                      (unsigned long long)p, (unsigned long long)a, (unsigned long long)b);
                ++products;
            }
        for (int k = 0; k < 8; ++k) {
            const uint64_t a = rnd_below(p), b = rnd_below(p), e1 = rnd(), e2 = rnd_below(1ull << 40);
            const uint64_t am = f.from(a, f.r2()), bm = f.from(b, f.r2());
            CHECK(am == mulmod(a, R, p), "from p=%llu", (unsigned long long)p);
            CHECK(fpow(f, am, e1) == mulmod(powmod(a, e1, p), R, p), "fpow p=%llu", (unsigned long long)p);
            uint64_t r1, r2;
            fpow2(f, am, bm, e2, r1, r2);
            CHECK(r1 == mulmod(powmod(a, e2, p), R, p) && r2 == mulmod(powmod(b, e2, p), R, p), "fpow2");
            fpow_2e(f, am, e1, e2, r1, r2);
            CHECK(r1 == mulmod(powmod(a, e1, p), R, p) && r2 == mulmod(powmod(a, e2, p), R, p), "fpow_2e");
            CHECK(f.dbl(am) == mulmod(2 * (u128)a % p, R, p), "dbl");
        }
        if (p > 37) CHECK(fermat2(f, p) == (powmod(2, p - 1, p) == 1), "fermat2 mont p=%llu", (unsigned long long)p);
    }
    // long dependent chains: every product's output is the next input (x <- x * y)
    for (int c = 0; c < 20 * scale; ++c) {
        const uint64_t p = (rnd_below(1ull << 62) | 1) | (1ull << 61);
        if (p >= (1ull << 62)) continue;
        MontField f;
        f.init(p);
        uint64_t x = rnd_below(p), y = rnd_below(p), ref = x;
        const uint64_t Rinv_y = f.mul(y, 1);  // y * 2^-64
        for (int k = 0; k < 2000; ++k) {
            x = f.mul(x, y);
            ref = mulmod(ref, Rinv_y, p);
        }
        CHECK(x == ref, "mont chain p=%llu", (unsigned long long)p);
        products += 4000;
    }
    std::printf("montgomery: %zu moduli, %ld products\n", mods.size(), products);
}

// ------------------------------------------------------------------ FP64 field, p < 2^50
static long long as_int(double x) { return (long long)x; }
static uint64_t canon(long long v, uint64_t p) {
    const long long r = v % (long long)p;
    return (uint64_t)(r < 0 ? r + (long long)p : r);
}

static void check_f64(int scale) {
    std::vector<uint64_t> mods = {3, 5, 7, 65537, 4294967291ull, (1ull << 50) - 27, (1ull << 50) - 1,
                                  (1ull << 49) + 1, 1000000000039ull, 95000000000011ull};
    for (int i = 0; i < 200 * scale; ++i) mods.push_back((rnd_below(1ull << (2 + rnd_below(48))) | 1) | 3);
    long products = 0, hashes = 0;
    for (uint64_t p : mods) {
        if (p >= (1ull << 50)) continue;
        F64Field f;
        f.init(p);
        const double P = (double)p;
        std::vector<double> ops = {0.0, -0.0, 1.0, -1.0, P - 1, -(P - 1), std::floor(P / 2), -std::floor(P / 2)};
        for (int k = 0; k < 48; ++k) {
            const double v = (double)rnd_below(p);
            ops.push_back((rnd() & 1) ? v : -v);
        }
        for (double a : ops) {
            // the canonical residue and the hash word
            const uint64_t ca = canon(as_int(a), p);
            CHECK(f.key(a) == ca, "key p=%llu a=%.0f", (unsigned long long)p, a);
            CHECK(f.hash_word(a) == (uint32_t)ca * F64Field::kHashMul, "hash p=%llu a=%.0f", (unsigned long long)p, a);
            ++hashes;
            for (double b : ops) {
                const double r = f.mul(a, b);
                CHECK(r == std::floor(r) && r > -P && r < P, "f64 range p=%llu", (unsigned long long)p);
                CHECK(canon(as_int(r), p) == mulmod(canon(as_int(a), p), canon(as_int(b), p), p),
                      "f64 mul p=%llu a=%.0f b=%.0f r=%.0f", (unsigned long long)p, a, b, r);
                ++products;
            }
            // a times small integers with |a * s| < 2^52
            for (int k = 0; k < 8; ++k) {
                const double lim = std::floor(4503599627370496.0 / std::max(1.0, std::fabs(a)));
                double s = (double)rnd_below((uint64_t)std::min(lim, 1e15));
                if (rnd() & 1) s = -s;
                const double r = f.mul_small(a, s);
                CHECK(r == std::floor(r) && r > -P && r < P, "mul_small range");
                CHECK(canon(as_int(r), p) == canon((long long)(canon(as_int(a), p) * (u128)canon(as_int(s), p) % p), p),
                      "mul_small p=%llu a=%.0f s=%.0f", (unsigned long long)p, a, s);
            }
        }
        // dependent chains of products (inputs from the signed output range)
        double x = (double)rnd_below(p), y = -(double)rnd_below(p);
        uint64_t ref = canon(as_int(x), p);
        const uint64_t yc = canon(as_int(y), p);
        for (int k = 0; k < 1000 * scale; ++k) {
            x = f.mul(x, y);
            ref = mulmod(ref, yc, p);
        }
        CHECK(f.key(x) == ref, "f64 chain p=%llu", (unsigned long long)p);
        products += 1000 * scale;
        // powers, the paired powers and the small-integer powers (b and -c as small integers)
        for (int k = 0; k < 6; ++k) {
            const uint64_t a = rnd_below(p), b = rnd_below(p), e1 = rnd_below(p), e2 = rnd();
            CHECK(f.key(fpow(f, (double)a, e2)) == powmod(a, e2, p), "f64 fpow");
            double r1, r2;
            fpow2(f, (double)a, (double)b, e1, r1, r2);
            CHECK(f.key(r1) == powmod(a, e1, p) && f.key(r2) == powmod(b, e1, p), "f64 fpow2");
            fpow_2e(f, (double)a, e1, e2, r1, r2);
            CHECK(f.key(r1) == powmod(a, e1, p) && f.key(r2) == powmod(a, e2, p), "f64 fpow_2e");
            const uint64_t smax = (1ull << 52) / p;
            if (smax >= 2) {
                const uint64_t sa = 1 + rnd_below(smax - 1), sb = 1 + rnd_below(smax - 1);
                fpow2_small(f, (double)sa, -(double)sb, e1, r1, r2);
                CHECK(f.key(r1) == powmod(sa % p, e1, p) && f.key(r2) == powmod(p - sb % p, e1, p),
                      "fpow2_small p=%llu sa=%llu sb=%llu", (unsigned long long)p, (unsigned long long)sa,
                      (unsigned long long)sb);
                CHECK(f.key(fpow_small(f, (double)sa, e1)) == powmod(sa % p, e1, p) &&
                          f.key(fpow_small(f, -(double)sb, e2)) == powmod(p - sb % p, e2, p) &&
                          f.key(fpow_small(f, -(double)sb, 0)) == 1 % p,
                      "fpow_small p=%llu sa=%llu sb=%llu", (unsigned long long)p, (unsigned long long)sa,
                      (unsigned long long)sb);
            }
        }
        if (p > 37) CHECK(fermat2(f, p) == (powmod(2, p - 1, p) == 1), "fermat2 f64 p=%llu", (unsigned long long)p);
    }
    std::printf("fp64: %zu moduli, %ld products, %ld hash words\n", mods.size(), products, hashes);
}

// ------------------------------------------------------------------ small-modulus remainders
static void check_small_mod(int scale) {
    long n = 0;
    for (uint32_t d = 1; d < 65536; ++d) {
        const uint32_t r32 = (uint32_t)((1ull << 32) % d);
        const double dinv = 1.0 / (double)d;
        uint32_t av[16] = {0, 1, d - 1, d, d + 1, 2 * d - 1, 0xFFFFFFFFu, 0xFFFFFFFFu - d, 0x80000000u, 65535u, 65536u};
        for (int k = 11; k < 16; ++k) av[k] = (uint32_t)rnd();
        for (uint32_t a : av) {
            CHECK(mod16_f(a, d, dinv) == a % d, "mod16_f a=%u d=%u", a, d);
            if (a < 65536) CHECK(mod16(a, d) == a % d, "mod16 a=%u d=%u", a, d);
            n += 2;
        }
        for (int k = 0; k < 4 * scale; ++k) {
            const uint32_t a = (uint32_t)rnd_below(65536);
            CHECK(mod16(a, d) == a % d, "mod16 a=%u d=%u", a, d);
            const uint64_t x = (k & 1) ? rnd() : ~0ull - rnd_below(1000);
            CHECK(mod_small(x, d, r32) == x % d && mod_small_f(x, d, r32, dinv) == x % d, "mod_small x=%llu d=%u",
                  (unsigned long long)x, d);
            n += 3;
        }
    }
    std::printf("small moduli: every divisor below 2^16, %ld remainders\n", n);
}

// ------------------------------------------------------------------ 52-bit helpers, exact division
static void check_wide(int scale) {
    long n = 0;
    const uint64_t Qs[] = {367567200ull, 367567200ull * 19, 367567200ull * 19 * 23, 367567200ull * 19 * 23 * 29,
                           367567200ull * 19 * 23 * 29 * 31, (1ull << 52) - 1, (1ull << 52) - 3, 2, 3};
    for (int i = 0; i < 2000 * scale; ++i) {
        const uint64_t m = 1 + rnd_below((1ull << (1 + rnd_below(52))) - 1);
        const uint64_t a = rnd_below(1ull << 52);
        uint64_t rem;
        const uint64_t q = divmod52(a, m, rem);
        CHECK(q == a / m && rem == a % m, "divmod52 a=%llu m=%llu", (unsigned long long)a, (unsigned long long)m);
        const uint64_t Q = (i % 3 == 0) ? Qs[rnd_below(9)] : 2 + rnd_below((1ull << 52) - 2);
        const uint64_t x = rnd_below(65536), y = rnd_below(Q);
        CHECK(mulmod_q52(x, y, Q, 1.0 / (double)Q) == (uint64_t)((u128)x * y % Q), "mulmod_q52 x=%llu y=%llu Q=%llu",
              (unsigned long long)x, (unsigned long long)y, (unsigned long long)Q);
        const uint64_t d = rnd() | 1, v = rnd_below(~0ull / d);
        CHECK(exact_div_odd(v * d, d) == v, "exact_div_odd");
        const uint64_t s = rnd() >> rnd_below(64);
        const uint64_t r = isqrt_ceil64(s);
        CHECK((u128)r * r >= s && (r == 0 || (u128)(r - 1) * (r - 1) < s), "isqrt_ceil64 %llu", (unsigned long long)s);
        CHECK(bitlen64(s) == (s ? 64 - (uint32_t)__builtin_clzll(s) : 0u), "bitlen64");
        n += 5;
    }
    for (uint64_t Q : Qs) {  // the edges of mulmod_q52's range
        const uint64_t xs[3] = {0, 1, 65535}, ys[4] = {0, 1, Q - 1, Q / 2};
        for (uint64_t x : xs)
            for (uint64_t y : ys) {
                CHECK(mulmod_q52(x, y, Q, 1.0 / (double)Q) == (uint64_t)((u128)x * y % Q), "mulmod_q52 edge");
                ++n;
            }
    }
    std::printf("52-bit helpers: %ld checks\n", n);
}

// ------------------------------------------------------------------ Jacobi symbols
static int euler(uint64_t a, uint64_t p) {  // (a/p) for an odd prime p
    const uint64_t e = powmod(a % p, (p - 1) / 2, p);
    return e == 0 ? 0 : e == 1 ? 1 : -1;
}

static void check_jacobi(int scale) {
    long n = 0;
    for (int i = 0; i < 300 * scale; ++i) {
        const uint64_t p = random_prime(3, (i & 1) ? (1ull << 32) : (1ull << 62));
        const uint64_t q = random_prime(3, 1ull << 31);
        for (int k = 0; k < 8; ++k) {
            const uint64_t a = (k < 2) ? (uint64_t)k : rnd_below(p);
            CHECK(jacobi64(a, p) == euler(a, p), "jacobi64 a=%llu p=%llu", (unsigned long long)a, (unsigned long long)p);
            if (p < (1ull << 32)) CHECK(jacobi32((uint32_t)a, (uint32_t)p) == euler(a, p), "jacobi32");
            const uint32_t s = (uint32_t)rnd_below(65536);
            CHECK(jacobi_small(s, p) == euler(s, p), "jacobi_small s=%u p=%llu", s, (unsigned long long)p);
            const int64_t x = (int64_t)(rnd() >> 1) * ((rnd() & 1) ? -1 : 1) / (int64_t)(1 + rnd_below(1ull << 40));
            const uint64_t xm = x < 0 ? (p - (uint64_t)(-(x + 0)) % p) % p : (uint64_t)x % p;
            CHECK(jacobi_signed(x, p) == euler(xm, p), "jacobi_signed x=%lld p=%llu", (long long)x, (unsigned long long)p);
            // a composite modulus: the symbol is multiplicative in the modulus
            if (p < (1ull << 31) && p != q) {
                const uint64_t pq = p * q, b = rnd_below(pq);
                CHECK(jacobi64(b, pq) == euler(b, p) * euler(b, q), "jacobi64 composite");
            }
            n += 5;
        }
    }
    std::printf("jacobi: %ld symbols\n", n);
}

// The BSGS bucket tests: bucket_needs_look and bucket_may_hold (packed 16-bit minimum over
// w ^ fp_pair(hw)) against slot-by-slot compares, on buckets with 0 to 8 filled slots, for lookups of stored, absent,
// 0xFFFE and raw 0xFFFF fingerprints under random high halves of the hash word; and the property
// the kernel relies on: a lookup whose clamped fingerprint is stored in the bucket, or whose bucket
// is full, always looks (so no stored key is missed, and a raw 0xFFFF always looks).
static void check_bucket(int scale) {
    const long n = 400000L * scale;
    for (long it = 0; it < n; ++it) {
        uint16_t sl[8];
        const int fill = (int)rnd_below(9);
        for (int k = 0; k < 8; ++k) {
            uint16_t v = 0xFFFF;
            if (k < fill) {
                const uint64_t r = rnd_below(8);
                v = r == 0 ? 0xFFFE : r == 1 ? 0 : (uint16_t)umin((uint32_t)(rnd() & 0xFFFF), 0xFFFEu);
            }
            sl[k] = v;
        }
        uint4 w;
        w.x = sl[0] | (uint32_t)sl[1] << 16;
        w.y = sl[2] | (uint32_t)sl[3] << 16;
        w.z = sl[4] | (uint32_t)sl[5] << 16;
        w.w = sl[6] | (uint32_t)sl[7] << 16;
        const uint64_t pick = rnd_below(6);
        uint32_t lo = pick == 0 ? 0xFFFF : pick == 1 ? 0xFFFE : pick == 2 && fill ? sl[rnd_below(fill)] : rnd() & 0xFFFF;
        const uint32_t hw = (uint32_t)(rnd() & 0xFFFF0000u) | lo;
        CHECK(fp_pair(hw) == (lo | lo << 16), "fp_pair %08x", hw);
        bool any = false;
        uint32_t mm = 0;
        for (int k = 0; k < 8; ++k) {
            any |= sl[k] == lo;
            if (sl[k] == lo) mm |= 1u << k;
        }
        const bool full = sl[7] != 0xFFFF;
        const bool look = bucket_needs_look(w, fp_pair(hw));
        CHECK(look == (any || full), "needs_look fill %d lo %04x: %d", fill, lo, (int)look);
        bool near = false;  // a slot holds lo or lo ^ 1
        for (int k = 0; k < 8; ++k) near |= (sl[k] ^ lo) <= 1u;
        const bool may = bucket_may_hold(w, fp_pair(hw));
        CHECK(may == near, "may_hold fill %d lo %04x: %d", fill, lo, (int)may);
        CHECK(match_mask(w, lo) == mm, "match_mask fill %d lo %04x", fill, lo);
        const uint32_t clamped = umin(lo, 0xFFFEu);
        bool stored = false;
        for (int k = 0; k < fill; ++k) stored |= sl[k] == clamped;
        if (stored || full || lo == 0xFFFF) CHECK(look, "missed lookup fill %d lo %04x", fill, lo);
        // a table without overflow: a stored key is always found in its own bucket, also for a raw
        // 0xFFFF (stored as 0xFFFE)
        if (stored) CHECK(may, "missed lookup without overflow fill %d lo %04x", fill, lo);
    }
    std::printf("bucket: %ld lookups on 0-8 filled slots\n", n);
}

// ------------------------------------------------------------------ medium primes of p - 1
// The table of the order-mode setup (make_ext_table), the exact divisibility test and quotient by
// r^-1 mod 2^64, inv_mod_prime for every residue, and narrow_class against the CRT in 128 bits.
static void check_ext(int scale) {
    static constexpr ExtTable T = make_ext_table();
    long n = 0;
    uint32_t prev = 17;
    for (uint32_t e = 0; e < kExtPrimes; ++e) {
        const ExtPrime& P = T.e[e];
        CHECK(P.r > prev && is_prime(P.r), "ext table: %u after %u", P.r, prev);
        for (uint32_t q = prev + 2; q < P.r; q += 2) CHECK(!is_prime(q), "ext table skips the prime %u", q);
        prev = P.r;
        CHECK(P.rinv * P.r == 1 && P.lim == ~0ull / P.r && P.r32 == (uint32_t)((1ull << 32) % P.r) &&
                  P.dinv == 1.0 / (double)P.r,
              "ext constants r=%u", P.r);
        for (int k = 0; k < 200 * scale; ++k) {
            const uint64_t x = (k & 1) ? rnd() : P.r * rnd_below(~0ull / P.r + 1);
            const bool div = x * P.rinv <= P.lim;
            CHECK(div == (x % P.r == 0), "ext divisibility x=%llu r=%u", (unsigned long long)x, P.r);
            if (div) CHECK(x * P.rinv == x / P.r, "ext exact quotient");
            n += 2;
        }
        for (uint32_t a = 1; a < P.r; ++a) {
            CHECK(a * inv_mod_prime(a, P.r, P.dinv) % P.r == 1, "inv_mod_prime a=%u r=%u", a, P.r);
            ++n;
        }
        for (int k = 0; k < 100 * scale; ++k) {
            uint64_t L = 1 + rnd_below(1ull << (1 + rnd_below(52)));
            if (k == 0) L = (1ull << 62) / P.r - 1;  // the largest L * r the setup allows
            const uint64_t nst = rnd_below(1ull << 62);
            const uint32_t kr = (uint32_t)rnd_below(P.r);
            uint64_t L2 = L, n2 = nst;
            const bool ch = narrow_class(L2, n2, kr, P.r, P.r32, P.dinv);
            if (L % P.r == 0) {
                CHECK(!ch && L2 == L && n2 == nst, "narrow_class changed a class r | L");
            } else {
                CHECK(ch && L2 == L * P.r && n2 >= nst && (n2 - nst) % L == 0 && (u128)(n2 - nst) < (u128)L * P.r &&
                          n2 % P.r == kr,
                      "narrow_class L=%llu nst=%llu k=%u r=%u", (unsigned long long)L, (unsigned long long)nst, kr, P.r);
            }
            ++n;
        }
    }
    std::printf("medium primes: %u primes %u..%u, %ld checks\n", kExtPrimes, T.e[0].r, prev, n);
}

// Small bases in the Montgomery field: mont_mul_small against 128-bit a * s mod p (p odd up to
// 2^62, s up to 2^24 - 1, a at 0, p - 1 and random), and mont_pow2_small / mont_pow_small against
// fpow of the Montgomery form (right to left, full products) for windows of 1 to 4 bits with the
// largest bases each width allows (the item constants' rule: s^(2^w - 1) < 2^24), exponents 0, 1,
// 2^k and random up to 2^62; mont_neg against p - x.
static void check_mont_small(int scale) {
    const int n_mod = 300 * scale;
    long n = 0;
    for (int i = 0; i < n_mod; ++i) {
        const uint64_t p = (i % 4 == 0) ? ((1ull << 62) - 57) : ((rnd_below(1ull << (51 + rnd_below(11))) | 1) | (1ull << 50));
        MontField f;
        f.init(p);
        const double pinv = 1.0 / (double)p;
        for (int k = 0; k < 64; ++k) {
            const uint64_t a = k == 0 ? 0 : k == 1 ? p - 1 : rnd_below(p);
            const uint32_t s = k == 2 ? (1u << 24) - 1 : k == 3 ? 1u : (uint32_t)(1 + rnd_below((1u << (1 + rnd_below(24))) - 1));
            const uint64_t ref = (uint64_t)((u128)a * s % p);
            CHECK(mont_mul_small(a, s, p, pinv) == ref, "mont_mul_small p %llu a %llu s %u", (unsigned long long)p,
                  (unsigned long long)a, s);
            CHECK(mont_neg(f, a) == (a ? p - a : 0), "mont_neg");
            ++n;
        }
        static const uint32_t smax[5] = {0, (1u << 24) - 1, 255, 10, 3};  // largest |s| per window width
        for (uint32_t w = 1; w <= 4; ++w) {
            for (int k = 0; k < 6; ++k) {
                const uint32_t sa = k == 0 ? smax[w] : 1 + (uint32_t)rnd_below(smax[w]);
                const uint32_t sb = k == 1 ? 1u : 1 + (uint32_t)rnd_below(smax[w]);
                uint32_t pa[4] = {0, 0, 0, 0}, pb[4] = {0, 0, 0, 0};
                for (uint32_t j = 0; j < w; ++j) {
                    uint64_t xa = sa, xb = sb;
                    for (uint32_t q = 0; q < j; ++q) {
                        xa *= xa;
                        xb *= xb;
                    }
                    pa[j] = (uint32_t)xa;
                    pb[j] = (uint32_t)xb;
                }
                const uint64_t e = k == 2 ? 0 : k == 3 ? 1 : k == 4 ? 1ull << rnd_below(62) : rnd_below(1ull << 62);
                const uint64_t ma = f.from(sa % p, f.r2()), mb = f.from(sb % p, f.r2());
                const uint64_t ra = fpow(f, ma, e), rb = fpow(f, mb, e);
                uint64_t ga, gb;
                mont_pow2_small(f, pa, pb, w, e, pinv, ga, gb);
                CHECK(ga == ra && gb == rb, "mont_pow2_small p %llu w %u s %u %u e %llu", (unsigned long long)p, w, sa,
                      sb, (unsigned long long)e);
                CHECK(mont_pow_small(f, pa, w, e, pinv) == ra, "mont_pow_small p %llu w %u s %u e %llu",
                      (unsigned long long)p, w, sa, (unsigned long long)e);
                ++n;
            }
        }
    }
    std::printf("mont small bases: %ld cases\n", n);
}

int main(int argc, char** argv) {
    const int scale = argc > 1 ? std::atoi(argv[1]) : 1;
    check_mont(scale);
    check_f64(scale);
    check_small_mod(scale);
    check_wide(scale);
    check_jacobi(scale);
    check_bucket(scale);
    check_ext(scale);
    check_mont_small(scale);
    std::printf("ok: %ld checks\n", g_checks);
    return 0;
}
