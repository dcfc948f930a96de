// Device arithmetic for the MI355X (gfx950) sieve worker.
//
// Two exact prime fields share one interface so that every stage is a template over the field:
//   MontField : 64-bit Montgomery arithmetic, odd p < 2^62, R = 2^64.  The REDC carry out of the
//               low word is (lo != 0), because lo + (lo * ninv mod 2^64) * p == 0 mod 2^64, so the
//               low half of m*p is never formed: 11 32-bit multiplies per product.
//   F64Field  : exact modular products in IEEE double for p < 2^50, on the full-rate FP64 FMA
//               pipes of CDNA4.  x = a*b is split exactly as h + l (h = fl(a*b), l = fma(a,b,-h));
//               the quotient q = rint(h/p) comes from one FMA against 1.5*2^52; r = fma(-q,p,h) + l
//               is exact because |h - q*p| < 2^53.  |q - x/p| < 0.76, so r lies in (-p, p): values
//               are kept in that signed range and made canonical only where they are compared.
// Both give the canonical residue through key(), which every hash and equality test uses.
//
// Every function here also compiles for the host (MPS_HD = __host__ __device__), where it runs on
// the CPU's 64-bit integers and IEEE doubles (fma from the C library, which rounds once, as
// v_fma_f64 does): the CPU tests check the same code against 128-bit and exact arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MPS_HD __host__ __device__ __forceinline__

namespace mps {

MPS_HD uint64_t dbits(double x) { return __builtin_bit_cast(uint64_t, x); }
MPS_HD double dfrom(uint64_t b) { return __builtin_bit_cast(double, b); }
MPS_HD double fma_rn(double a, double b, double c) { return __builtin_fma(a, b, c); }
MPS_HD uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// x mod d for d < 2^16, from 32-bit remainders only; r32 = 2^32 mod d.
MPS_HD uint32_t mod_small(uint64_t x, uint32_t d, uint32_t r32) {
    const uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
    const uint32_t r = (hi % d) * r32 + (lo % d);
    return r % d;
}

// a mod d for d < 2^16 with dinv = 1.0 / d (rounded to nearest, from the host).  a and d are exact
// doubles and a * dinv lies within 2^-20 of a / d, which is an integer or at least 1/d >= 2^-16
// below one, so the truncated quotient is floor(a / d) or one less: r < 2d, and min corrects it.
MPS_HD uint32_t mod16_f(uint32_t a, uint32_t d, double dinv) {
    const uint32_t q = (uint32_t)((double)a * dinv);
    const uint32_t r = a - q * d;
    return umin(r, r - d);
}

// mod_small for a divisor fixed per item, with its reciprocal: three quotients from FP64 products
// in place of three integer divisions ((d - 1) * r32 + d - 1 < 2^32 for d <= 2^16).
MPS_HD uint32_t mod_small_f(uint64_t x, uint32_t d, uint32_t r32, double dinv) {
    const uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
    return mod16_f(mod16_f(hi, d, dinv) * r32 + mod16_f(lo, d, dinv), d, dinv);
}

MPS_HD uint32_t gcd32(uint32_t a, uint32_t b) {
    while (b) {
        const uint32_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// Jacobi symbol (a/n) for odd n < 2^32 and a < n (callers pass a reduced a).
MPS_HD int jacobi32(uint32_t a, uint32_t n) {
    int s = 1;
    while (a != 0) {
        const uint32_t e = __builtin_ctz(a);
        a >>= e;
        if ((e & 1) && ((n & 7) == 3 || (n & 7) == 5)) s = -s;
        if ((a & 3) == 3 && (n & 3) == 3) s = -s;
        const uint32_t r = n % a;
        n = a;
        a = r;
    }
    return n == 1 ? s : 0;
}

// Jacobi symbol (a/n) for odd n < 2^63, binary method (no 64-bit division).
MPS_HD int jacobi64(uint64_t a, uint64_t n) {
    int s = 1;
    while (a != 0) {
        const uint32_t e = (uint32_t)__builtin_ctzll(a);
        a >>= e;
        if ((e & 1) && ((n & 7) == 3 || (n & 7) == 5)) s = -s;
        if (a < n) {
            const uint64_t x = a;
            a = n;
            n = x;
            if ((a & 3) == 3 && (n & 3) == 3) s = -s;
        }
        a -= n;
    }
    return n == 1 ? s : 0;
}

// (a/p) for a < 2^16, p odd < 2^63: strip the twos, then one reciprocity step brings the
// problem into 32 bits.
MPS_HD int jacobi_small(uint32_t a, uint64_t p) {
    if (a == 0) return 0;
    int s = 1;
    const uint32_t e = __builtin_ctz(a);
    a >>= e;
    const uint32_t p8 = (uint32_t)p & 7;
    if ((e & 1) && (p8 == 3 || p8 == 5)) s = -s;
    if (a == 1) return s;
    if ((a & 3) == 3 && (p & 3) == 3) s = -s;
    const uint32_t r32 = (uint32_t)((0x100000000ULL) % a);
    return s * jacobi32(mod_small(p, a, r32), a);
}

// Jacobi symbol (x mod p / p) for x given as a signed 64-bit value (b or -c of an item).
MPS_HD int jacobi_signed(int64_t x, uint64_t p) {
    const uint64_t ax = x < 0 ? 0ull - (uint64_t)x : (uint64_t)x;  // 2^63 for INT64_MIN
    int s = 1;
    if (x < 0 && (p & 3) == 3) s = -1;  // (-1/p)
    if (ax < 65536u) return s * jacobi_small((uint32_t)ax, p);
    return s * jacobi64(ax % p, p);
}


// ---------------------------------------------------------------------------- integer helpers

MPS_HD uint64_t isqrt_ceil64(uint64_t x) {
    if (x == 0) return 0;
    uint64_t r = (uint64_t)sqrt((double)x);
    while (r > 0 && r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return r * r < x ? r + 1 : r;
}

MPS_HD uint32_t bitlen64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// Quotient of an exact division x / d (d odd): x times the inverse of d modulo 2^64.
MPS_HD uint64_t exact_div_odd(uint64_t x, uint64_t d) {
    uint64_t inv = d;  // d * d == 1 (mod 8): correct to 3 bits, doubled by each Newton step
    for (int k = 0; k < 5; ++k) inv *= 2 - d * inv;
    return x * inv;
}

// a mod m for a, m below 2^16 (m >= 1): the float quotient is within one of the true one.
MPS_HD uint32_t mod16(uint32_t a, uint32_t m) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float rm = __frcp_rn((float)m);
#else
    const float rm = 1.0f / (float)m;  // the same correctly rounded reciprocal
#endif
    const uint32_t q = (uint32_t)((float)a * rm);
    int32_t r = (int32_t)(a - q * m);
    if (r < 0) r += (int32_t)m;
    if (r >= (int32_t)m) r -= (int32_t)m;
    return (uint32_t)r;
}

// The inverse of a modulo a prime r < 2^16 (a not a multiple of r), as a^(r-2) mod r: products
// below r^2 < 2^32, each reduced by mod16_f with dinv = 1.0 / r.
MPS_HD uint32_t inv_mod_prime(uint32_t a, uint32_t r, double dinv) {
    uint32_t e = r - 2, x = a, y = 1;
    while (e) {
        if (e & 1) y = mod16_f(y * x, r, dinv);
        e >>= 1;
        if (e) x = mod16_f(x * x, r, dinv);
    }
    return y;
}

// Narrows the exponent class n == nst (mod L) by n == k (mod r), r a prime below 2^16 (r32 = 2^32
// mod r, dinv = 1.0 / r): the admissible n are n == nst + L * j (mod L * r) with
// j = (k - nst) * L^-1 mod r, and nst + L * j is the first of them from nst on.  False (nothing
// changed) when r divides L.  The caller keeps L * r and nst + L * (r - 1) below 2^64.
MPS_HD bool narrow_class(uint64_t& L, uint64_t& nst, uint32_t k, uint32_t r, uint32_t r32, double dinv) {
    const uint32_t Lr = mod_small_f(L, r, r32, dinv);
    if (Lr == 0) return false;
    const uint32_t nr = mod_small_f(nst, r, r32, dinv);
    const uint32_t d = k >= nr ? k - nr : k + r - nr;
    const uint32_t j = mod16_f(d * inv_mod_prime(Lr, r, dinv), r, dinv);
    nst += L * j;
    L *= r;
    return true;
}

// a mod m and a / m for a and m below 2^52 (m >= 1): both are exact doubles, so the double quotient
// is within one of the true one.
MPS_HD uint64_t divmod52(uint64_t a, uint64_t m, uint64_t& rem) {
    uint64_t q = (uint64_t)((double)a / (double)m);
    int64_t r = (int64_t)(a - q * m);
    if (r < 0) {
        r += m;
        --q;
    } else if (r >= (int64_t)m) {
        r -= m;
        ++q;
    }
    rem = (uint64_t)r;
    return q;
}

// a * b mod Qo for a < 2^16 and b < Qo < 2^52: the quotient from doubles (a, b exact; the product
// and the scaling by 1/Qo round) is within one of the true one, and a*b - q*Qo is exact modulo 2^64.
MPS_HD uint64_t mulmod_q52(uint64_t a, uint64_t b, uint64_t Qo, double Qo_inv) {
    const uint64_t q = (uint64_t)((double)a * (double)b * Qo_inv);
    int64_t r = (int64_t)(a * b - q * Qo);
    if (r < 0) r += (int64_t)Qo;
    if (r >= (int64_t)Qo) r -= (int64_t)Qo;
    return (uint64_t)r;
}

// ---------------------------------------------------------------------------------------------
struct MontField {
    typedef uint64_t T;
    static constexpr bool kKeepsOne = true;  // set() takes the Montgomery form of 1
    static constexpr bool kSmallMul = false;  // no mul_small (operands are in Montgomery form)
    static constexpr int kOrderSetupWaves = 6;  // order-mode setup: its LDS allows 6 workgroups per CU
    static constexpr uint32_t kGiantGroup = 1;  // k_bsgs: giant steps per branch on the lookups
    // k_bsgs: no second giant loop for tables without overflow (c4 BSGS 25.0 -> 25.6 ms with one)
    static constexpr bool kTightLookup = false;
    uint64_t p, ninv, one_;

    MPS_HD void init(uint64_t mod) {
        p = mod;
        uint64_t inv = (3 * mod) ^ 2;  // correct to 5 bits for odd mod
        for (int i = 0; i < 4; ++i) inv *= 2 - mod * inv;
        ninv = 0 - inv;
        one_ = (0 - mod) % mod;
    }
    MPS_HD void set(uint64_t mod, uint64_t ninv_, uint64_t one) {
        p = mod;
        ninv = ninv_;
        one_ = one;
    }
    MPS_HD T one() const { return one_; }
    // a * b * 2^-64 mod p, as a Montgomery product over 32-bit words (CIOS: each word of b adds
    // a * b_i, then m * p with m = t0 * (-p^-1) mod 2^32 clears the low word, which is shifted out).
    // With a, b < p < 2^62 every partial sum fits 64 bits: a0 * b_i + (the 64-bit accumulator) <
    // 2^62 + 2^63, so the second word takes the whole accumulator as the addend of one
    // v_mad_u64_u32 (its low word is the accumulator's low word plus a0 * b1, its high word carries
    // the accumulator's high word on); m * p0 + u0 < 2^64; and the result is below 2p, so one
    // subtraction, selected by its borrow, makes it canonical: the same value as the 64-bit REDC
    // (lo * ninv, high halves of a*b and m*p), in 10 32-bit multiplies.
    MPS_HD T mul(T a, T b) const {
        const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
        const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
        const uint32_t p0 = (uint32_t)p, p1 = (uint32_t)(p >> 32), n0 = (uint32_t)ninv;
        // word 0 of b: t = (a * b0 + m * p) / 2^32
        uint64_t u = (uint64_t)a0 * b0;
        uint32_t m = (uint32_t)u * n0;
        uint64_t v = (uint64_t)a1 * b0 + (u >> 32);
        uint64_t w = (uint64_t)m * p0 + (uint32_t)u;  // low word 0
        uint64_t t = (uint64_t)m * p1 + v + (w >> 32);
        // word 1: t = (t + a * b1 + m * p) / 2^32
        u = (uint64_t)a0 * b1 + t;
        m = (uint32_t)u * n0;
        v = (uint64_t)a1 * b1 + (u >> 32);
        w = (uint64_t)m * p0 + (uint32_t)u;
        t = (uint64_t)m * p1 + v + (w >> 32);
        uint64_t d;
        return __builtin_sub_overflow(t, p, &d) ? t : d;
    }
    MPS_HD T dbl(T a) const {
        const uint64_t r = a + a;
        return r >= p ? r - p : r;
    }
    // R^2 mod p as the Montgomery form of 2^64: square mont(2) six times.
    MPS_HD T r2() const {
        T x = dbl(one_);
        for (int i = 0; i < 6; ++i) x = mul(x, x);
        return x;
    }
    MPS_HD T from(uint64_t x, T rr) const { return mul(x, rr); }  // x < p
    MPS_HD uint64_t key(T a) const { return a; }
    // A 32-bit word of key(a) for hash tables: Montgomery forms a * 2^64 mod p are already spread
    // over all bits, so the low word serves.
    MPS_HD uint32_t hash_word(T a) const { return (uint32_t)a; }

    MPS_HD uint64_t aux() const { return ninv; }
    MPS_HD static uint64_t to_bits(T a) { return a; }
    MPS_HD static T from_bits(uint64_t a) { return a; }
};

struct F64Field {
    typedef double T;
    static constexpr bool kKeepsOne = false;  // set() ignores its third argument
    static constexpr bool kSmallMul = true;   // mul_small: a times a small integer
    static constexpr int kOrderSetupWaves = 7;  // order-mode setup: 72 registers, 7 waves per SIMD
    static constexpr uint32_t kGiantGroup = 2;  // k_bsgs: giant steps per branch on the lookups
    // k_bsgs: a second giant loop for tables without overflow (c2 BSGS 1.99 -> 1.91 ms)
    static constexpr bool kTightLookup = true;
    static constexpr uint32_t kHashMul = 0x9E3779B1u;
    double p, pinv;
    uint32_t plo;  // low word of p
    uint32_t plc;  // plo * kHashMul (mod 2^32)

    MPS_HD void init(uint64_t mod) {
        p = (double)mod;
        pinv = 1.0 / p;
        plo = (uint32_t)mod;
        plc = plo * kHashMul;
    }
    MPS_HD void set(uint64_t mod, uint64_t pinv_bits, uint64_t) {
        p = (double)mod;
        pinv = dfrom(pinv_bits);
        plo = (uint32_t)mod;
        plc = plo * kHashMul;
    }
    MPS_HD T one() const { return 1.0; }
    MPS_HD T mul(T a, T b) const {
        const double h = a * b;
        const double l = fma_rn(a, b, -h);
        const double q = fma_rn(h, pinv, 6755399441055744.0) - 6755399441055744.0;
        return fma_rn(-q, p, h) + l;
    }
    // a * s for an integer s with |a * s| < 2^52: the product is exact, so one rounding step
    // (q = round(x / p), x - q * p, exact in the FMA) reduces it; 4 operations instead of 6.
    MPS_HD T mul_small(T a, double s) const {
        const double x = a * s;
        const double q = fma_rn(x, pinv, 6755399441055744.0) - 6755399441055744.0;
        return fma_rn(-q, p, x);
    }
    MPS_HD T dbl(T a) const {
        double r = a + a;
        r = r >= p ? r - p : r;
        return r <= -p ? r + p : r;
    }
    MPS_HD T r2() const { return 1.0; }
    MPS_HD T from(uint64_t x, T) const { return (double)x; }
    // canonical residue as an integer: adding 2^52 puts the integer in the mantissa bits
    MPS_HD uint64_t key(T a) const {
        const double c = a < 0.0 ? a + p : a;
        return dbits(c + 4503599627370496.0) & 0x000FFFFFFFFFFFFFULL;
    }
    // A 32-bit word of key(a) for hash tables: the low word of key(a) times an odd constant (a
    // bijection of the low word, so equal words mean equal low words).  Plain residues have structure
    // (t * B^i below p are small multiples of powers of b), which the product spreads over all bits.
    // a + 1.5 * 2^52 (|a| < 2^51) holds a + 2^51 in its mantissa, whose low word is a mod 2^32; a
    // negative a stands for a + p.  The sign comes from the same bits: mantissa bit 51 (bit 19 of
    // the high word) is set exactly when a >= 0 (-0.0 included), and (lo + plo) * kHashMul is
    // lo * kHashMul + plc, so the correction is a select of a per-modulus constant: v_bfe_i32
    // spreads bit 51 over a word and v_bfi_b32 keeps plc where it is clear (two VALU; written
    // out because the compiler turns the same expression into a compare, a move and a select).
    // The modulus must be the same in every lane of the wavefront (k_bsgs, one entry per
    // workgroup, is the only caller): plc is an SGPR operand of the select.
    MPS_HD uint32_t hash_word(T a) const {
        const uint64_t x = dbits(a + 6755399441055744.0);
        uint32_t corr;  // plc when a < 0, else 0
#if defined(__HIP_DEVICE_COMPILE__)
        asm("v_bfe_i32 %0, %1, 19, 1\n\tv_bfi_b32 %0, %0, 0, %2" : "=&v"(corr) : "v"((uint32_t)(x >> 32)), "s"(plc));
#else
        corr = ((x >> 51) & 1) ? 0u : plc;  // the two instructions' result, for the CPU tests
#endif
        return (uint32_t)x * kHashMul + corr;
    }
    MPS_HD uint64_t aux() const { return dbits(pinv); }
    MPS_HD static uint64_t to_bits(T a) { return dbits(a); }
    MPS_HD static T from_bits(uint64_t a) { return dfrom(a); }
};

template <class F>
MPS_HD typename F::T fpow(const F& f, typename F::T base, uint64_t e) {
    typename F::T r = f.one();
    while (e) {
        if (e & 1) r = f.mul(r, base);
        e >>= 1;
        if (e) base = f.mul(base, base);
    }
    return r;
}

// x^e1 and x^e2 in one loop: the squarings of x are shared, the two products are independent.
template <class F>
MPS_HD void fpow_2e(const F& f, typename F::T x, uint64_t e1, uint64_t e2, typename F::T& r1,
                                        typename F::T& r2) {
    r1 = f.one();
    r2 = f.one();
    while (e1 | e2) {
        if (e1 & 1) r1 = f.mul(r1, x);
        if (e2 & 1) r2 = f.mul(r2, x);
        e1 >>= 1;
        e2 >>= 1;
        if (e1 | e2) x = f.mul(x, x);
    }
}

// sa^e and sb^e for integers sa, sb with |sa| * p, |sb| * p < 2^52 (F64Field), left to right:
// each bit squares both and, when set, multiplies them by sa and sb with mul_small (the base
// stays a small integer, so that product is exact).  Any integer of the residue class will do
// for sa and sb (a negative -c stands for p - c).
template <class F>
MPS_HD void fpow2_small(const F& f, double sa, double sb, uint64_t e, double& ra, double& rb) {
    ra = 1.0;
    rb = 1.0;
    if (e == 0) return;
    ra = f.mul_small(ra, sa);  // the top bit
    rb = f.mul_small(rb, sb);
    for (int k = 62 - __builtin_clzll(e); k >= 0; --k) {
        ra = f.mul(ra, ra);
        rb = f.mul(rb, rb);
        if ((e >> k) & 1) {
            ra = f.mul_small(ra, sa);
            rb = f.mul_small(rb, sb);
        }
    }
}

// s^e for an integer s with |s| * p < 2^52 (F64Field), left to right with mul_small, as
// fpow2_small for one base.
template <class F>
MPS_HD double fpow_small(const F& f, double s, uint64_t e) {
    if (e == 0) return 1.0;
    double r = f.mul_small(1.0, s);
    for (int k = 62 - __builtin_clzll(e); k >= 0; --k) {
        r = f.mul(r, r);
        if ((e >> k) & 1) r = f.mul_small(r, s);
    }
    return r;
}

// a^e and b^e in one loop (the exponent's bits are walked once for both); E is uint32_t or
// uint64_t.
template <class F, class E>
MPS_HD void fpow2(const F& f, typename F::T a, typename F::T b, E e, typename F::T& ra,
                                      typename F::T& rb) {
    ra = f.one();
    rb = f.one();
    while (e) {
        if (e & 1) {
            ra = f.mul(ra, a);
            rb = f.mul(rb, b);
        }
        e >>= 1;
        if (e) {
            a = f.mul(a, a);
            b = f.mul(b, b);
        }
    }
}

// ---------------------------------------------------------------- small bases, Montgomery field
// b and t = -c of an item are small integers; their powers modulo p > 2^50 are the bulk of the
// order-mode setup.  Left to right, w bits at a time, a power of a small base s costs w squarings
// and one product by the small integer s^digit per window, against a squaring and a full product
// per bit right to left.

// a * s mod p for a < p < 2^62 and 1 <= s < 2^24, pinv = 1.0 / p (a Montgomery form times a plain
// integer is the Montgomery form of the product).  q = the double a * s * pinv truncated: its
// relative error is a few 2^-53 on a value below 2^24, so q is floor(a * s / p) or one off, and
// a * s - q * p, from the low 64 bits of both products, lies in (-p, 2p); one correction.
MPS_HD uint64_t mont_mul_small(uint64_t a, uint32_t s, uint64_t p, double pinv) {
    const uint32_t q = (uint32_t)((double)a * (double)s * pinv);
    uint64_t r = a * (uint64_t)s - (uint64_t)q * p;
    if ((int64_t)r < 0)
        r += p;
    else if (r >= p)
        r -= p;
    return r;
}

// s^d for a w-bit digit d from sp[i] = s^(2^i), i < w (the caller keeps s^(2^w - 1) < 2^24).
MPS_HD uint32_t small_digit_pow(const uint32_t* sp, uint32_t w, uint32_t d) {
    uint32_t r = 1;
    for (uint32_t i = 0; i < w; ++i) r *= ((d >> i) & 1u) ? sp[i] : 1u;
    return r;
}

// The Montgomery forms of sa^e and sb^e (sa, sb given by their window powers, w bits per window,
// 1 <= w <= 4), left to right: the first window's digit, then per window w squarings of each and
// one mont_mul_small each.  The two chains are independent (they overlap in the pipeline).
MPS_HD void mont_pow2_small(const MontField& f, const uint32_t* sa, const uint32_t* sb, uint32_t w, uint64_t e,
                            double pinv, uint64_t& ra, uint64_t& rb) {
    ra = f.one();
    rb = f.one();
    if (e == 0) return;
    const uint32_t mask = (1u << w) - 1u;
    int k = (int)((63u - (uint32_t)__builtin_clzll(e)) / w * w);
    for (bool first = true; k >= 0; k -= (int)w, first = false) {
        if (!first)
            for (uint32_t i = 0; i < w; ++i) {
                ra = f.mul(ra, ra);
                rb = f.mul(rb, rb);
            }
        const uint32_t d = (uint32_t)(e >> k) & mask;
        ra = mont_mul_small(ra, small_digit_pow(sa, w, d), f.p, pinv);
        rb = mont_mul_small(rb, small_digit_pow(sb, w, d), f.p, pinv);
    }
}

// mont_pow2_small for one base.
MPS_HD uint64_t mont_pow_small(const MontField& f, const uint32_t* s, uint32_t w, uint64_t e, double pinv) {
    uint64_t r = f.one();
    if (e == 0) return r;
    const uint32_t mask = (1u << w) - 1u;
    int k = (int)((63u - (uint32_t)__builtin_clzll(e)) / w * w);
    for (bool first = true; k >= 0; k -= (int)w, first = false) {
        if (!first)
            for (uint32_t i = 0; i < w; ++i) r = f.mul(r, r);
        r = mont_mul_small(r, small_digit_pow(s, w, (uint32_t)(e >> k) & mask), f.p, pinv);
    }
    return r;
}

// -x mod p (x < p).
MPS_HD uint64_t mont_neg(const MontField& f, uint64_t x) { return x ? f.p - x : 0; }

template <class F>
MPS_HD bool fequal(const F& f, typename F::T a, typename F::T b) {
    return f.key(a) == f.key(b);
}

// 2^(p-1) == 1 (mod p) by left-to-right square-and-double.
template <class F>
MPS_HD bool fermat2(const F& f, uint64_t p) {
    const uint64_t e = p - 1;
    int k = 63 - __builtin_clzll(e);
    typename F::T x = f.one();
    for (; k >= 0; --k) {
        x = f.mul(x, x);
        if ((e >> k) & 1) x = f.dbl(x);
    }
    return f.key(x) == f.key(f.one());
}

// ---------------------------------------------------------------- BSGS table buckets
// A bucket is 8 slots of 16-bit fingerprints in four words (slot 2k in the low half of word k),
// filled in order; 0xFFFF marks an empty slot, so a stored fingerprint is clamped to 0xFFFE, and
// slot 7 taken means the bucket is full and its keys may go on in the next bucket.

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// The low half of hash word hw in both halves of a word (one byte permute on the device).
MPS_HD uint32_t fp_pair(uint32_t hw) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(hw, hw, 0x01000100u);
#else
    return (hw & 0xFFFFu) * 0x10001u;
#endif
}

// A slot of the bucket holds the fingerprint that f2 carries in both halves, or the bucket is full
// (the chain goes on): a slot matches when its half of w ^ f2 is zero, so the packed minimum over
// the four words has a zero half.  Packed 16-bit operations and one compare stand in for eight
// compares and the seven scalar ORs that joined their lane masks.  The lookup may use the raw low
// half of the hash word: a raw 0xFFFF matches every empty slot, and a bucket without one is full,
// so such a lookup always goes on to the exact walk, which uses the clamped value.
MPS_HD bool bucket_needs_look(uint4 w, uint32_t f2) {
    auto x = [f2](uint32_t a) { return __builtin_bit_cast(u16x2, a ^ f2); };
    const u16x2 m = __builtin_elementwise_min(__builtin_elementwise_min(x(w.x), x(w.y)),
                                              __builtin_elementwise_min(x(w.z), x(w.w)));
    return (umin(m.x, m.y) == 0u) | (w.w < 0xFFFF0000u);
}

// The lookup test of a table in which no bucket overflowed (every key sits in its own bucket, so
// a full bucket ends the chain): a slot holds the raw fingerprint of f2 or that value with bit 0
// flipped.  The second form covers a raw 0xFFFF, whose key is stored clamped to 0xFFFE; for any
// other lookup it only adds rare false matches, which the exact walk drops.  The same instruction
// count as bucket_needs_look less its test of slot 7.
MPS_HD bool bucket_may_hold(uint4 w, uint32_t f2) {
    auto x = [f2](uint32_t a) { return __builtin_bit_cast(u16x2, a ^ f2); };
    const u16x2 m = __builtin_elementwise_min(__builtin_elementwise_min(x(w.x), x(w.y)),
                                              __builtin_elementwise_min(x(w.z), x(w.w)));
    return umin(m.x, m.y) <= 1u;
}

// Bit k set when slot k of the bucket holds fingerprint fp.
MPS_HD uint32_t match_mask(uint4 w, uint32_t fp) {
    return (((w.x & 0xFFFFu) == fp) ? 1u : 0u) | (((w.x >> 16) == fp) ? 2u : 0u) |
           (((w.y & 0xFFFFu) == fp) ? 4u : 0u) | (((w.y >> 16) == fp) ? 8u : 0u) |
           (((w.z & 0xFFFFu) == fp) ? 16u : 0u) | (((w.z >> 16) == fp) ? 32u : 0u) |
           (((w.w & 0xFFFFu) == fp) ? 64u : 0u) | (((w.w >> 16) == fp) ? 128u : 0u);
}

}  // namespace mps
