// Exact host checks: deterministic Miller-Rabin, the exact baby-step giant-step of run_bsgs and
// the multiplicative order (Math.hpp:209-310, :377-542).  They finish the numbers
// mps_gpu_sieve_filter returns (the reference's superset contract: base-2 Fermat pseudoprimes may
// be among them) and the records the device leaves to the host (M = MPS_HOST_DECIDES), and give
// q = ord_p(b) for every relation.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "mps_host.h"

namespace {

typedef unsigned __int128 u128;
typedef __int128 i128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)(a % m) * (b % m) % m); }

uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    if (m == 1) return 0;
    uint64_t r = 1;
    b %= m;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod(r, b, m);
        b = mulmod(b, b, m);
    }
    return r;
}

// Montgomery arithmetic, odd n < 2^63, R = 2^64.
struct Mont {
    uint64_t n, ninv, one, r2;
    explicit Mont(uint64_t mod) : n(mod) {
        uint64_t inv = (3 * mod) ^ 2;
        for (int i = 0; i < 4; ++i) inv *= 2 - mod * inv;
        ninv = 0 - inv;
        one = (0 - mod) % mod;
        r2 = (uint64_t)((u128)one * one % mod);
    }
    uint64_t mul(uint64_t a, uint64_t b) const {
        const u128 x = (u128)a * b;
        const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
        const uint64_t m = lo * ninv;
        const uint64_t t = hi + (uint64_t)(((u128)m * n) >> 64) + (lo != 0);
        return t >= n ? t - n : t;
    }
    uint64_t to(uint64_t a) const { return mul(a % n, r2); }
    uint64_t pow(uint64_t b, uint64_t e) const {
        uint64_t r = one;
        for (; e; e >>= 1) {
            if (e & 1) r = mul(r, b);
            b = mul(b, b);
        }
        return r;
    }
    int legendre(uint64_t am) const {
        if (am == 0) return 0;
        return pow(am, (n - 1) >> 1) == one ? 1 : -1;
    }
};

uint64_t gcd(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

uint64_t isqrt_ceil(uint64_t x) {
    if (x == 0) return 0;
    uint64_t r = (uint64_t)std::sqrt((long double)x);
    while (r > 0 && (u128)r * r > x) --r;
    while ((u128)(r + 1) * (r + 1) <= x) ++r;
    return (u128)r * r < x ? r + 1 : r;
}

// Pollard rho, Brent's variant with batched gcds; n odd composite.
uint64_t rho(uint64_t n, uint64_t seed) {
    uint64_t st = seed | 1;
    auto rnd = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    for (;;) {
        const uint64_t c = 1 + rnd() % (n - 1);
        auto f = [&](uint64_t x) { return (uint64_t)(((u128)x * x + c) % n); };
        uint64_t y = rnd() % n, x = y, ys = y, q = 1, g = 1;
        for (uint64_t r = 1; g == 1; r <<= 1) {
            x = y;
            for (uint64_t i = 0; i < r; ++i) y = f(y);
            for (uint64_t k = 0; k < r && g == 1; k += 128) {
                ys = y;
                const uint64_t lim = std::min<uint64_t>(128, r - k);
                for (uint64_t i = 0; i < lim; ++i) {
                    y = f(y);
                    q = mulmod(q, x > y ? x - y : y - x, n);
                }
                g = gcd(q, n);
            }
        }
        if (g == n) {
            do {
                ys = f(ys);
                g = gcd(x > ys ? x - ys : ys - x, n);
            } while (g == 1);
        }
        if (g != n) return g;
    }
}

void distinct_factors(uint64_t n, std::vector<uint64_t>& out) {
    out.clear();
    for (uint64_t d = 2; d < 200 && d * d <= n; ++d) {
        if (n % d) continue;
        out.push_back(d);
        while (n % d == 0) n /= d;
    }
    std::vector<uint64_t> stack;
    if (n > 1) stack.push_back(n);
    uint64_t seed = 0x9E3779B97F4A7C15ull ^ n;
    while (!stack.empty()) {
        const uint64_t x = stack.back();
        stack.pop_back();
        if (x == 1) continue;
        if (mps_is_prime64(x)) {
            out.push_back(x);
            continue;
        }
        const uint64_t d = rho(x, seed++);
        stack.push_back(d);
        stack.push_back(x / d);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

uint64_t mod_signed(int64_t v, uint64_t m) {
    const i128 r = (i128)v % (i128)m;
    return (uint64_t)(r < 0 ? r + (i128)m : r);
}

// Exact table for the baby steps: full 64-bit keys, first index kept.
struct Table {
    std::vector<uint64_t> keys;
    std::vector<uint32_t> vals;
    unsigned bits = 4;
    void reset(uint64_t entries) {
        bits = 4;
        while ((1ull << bits) < entries * 2) ++bits;
        if (keys.size() < (1ull << bits)) {
            keys.resize(1ull << bits);
            vals.resize(1ull << bits);
        }
        std::fill(keys.begin(), keys.begin() + (1ull << bits), ~0ull);
    }
    uint64_t slot(uint64_t k) const { return (k * 0xC2B2AE3D27D4EB4Full) >> (64 - bits); }
    void insert(uint64_t k, uint32_t v) {
        const uint64_t mask = (1ull << bits) - 1;
        for (uint64_t h = slot(k);; h = (h + 1) & mask) {
            if (keys[h] == k) return;
            if (keys[h] == ~0ull) {
                keys[h] = k;
                vals[h] = v;
                return;
            }
        }
    }
    bool find(uint64_t k, uint32_t& v) const {
        const uint64_t mask = (1ull << bits) - 1;
        for (uint64_t h = slot(k);; h = (h + 1) & mask) {
            if (keys[h] == ~0ull) return false;
            if (keys[h] == k) {
                v = vals[h];
                return true;
            }
        }
    }
};

thread_local Table tls_table;

// run_bsgs_reference (Math.hpp:488-542): n searched in [0, m^2) from n = 0.
int bsgs_plain(uint64_t prime, const mps_work_item* it, mps_relation* out) {
    if (prime < 2) return 0;
    const uint64_t bm = mod_signed(it->b, prime);
    const uint64_t t0 = mod_signed(it->c, prime);
    const uint64_t t = t0 == 0 ? 0 : prime - t0;
    if (bm == 0) {
        if (t == 0) {
            out->p = prime;
            out->M = (uint64_t)it->N0;
            out->q = 1;
            return 1;
        }
        return 0;
    }
    const uint64_t m = prime > (uint64_t)it->N1 ? isqrt_ceil((uint64_t)it->N1) : isqrt_ceil(prime);
    if (m == 0) return 0;
    Table& tab = tls_table;
    tab.reset(m);
    uint64_t pw = 1;
    for (uint64_t i = 0; i < m; ++i) {
        tab.insert(pw, (uint32_t)i);
        pw = mulmod(pw, bm, prime);
    }
    const uint64_t step_inv = powmod(powmod(bm, m, prime), prime - 2, prime);
    uint64_t giant = t;
    for (uint64_t j = 0; j < m; ++j) {
        uint32_t idx;
        if (tab.find(giant, idx)) {
            out->p = prime;
            out->M = idx + j * m;
            out->q = mps_multiplicative_order(bm, prime);
            return 1;
        }
        giant = mulmod(giant, step_inv, prime);
    }
    return 0;
}

}  // namespace

extern "C" {

int mps_is_prime64(uint64_t n) {
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (uint64_t b : bases) {
        if (n == b) return 1;
        if (n % b == 0) return 0;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++s;
    }
    for (uint64_t a : bases) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool witness = true;
        for (int r = 1; r < s && witness; ++r) {
            x = mulmod(x, x, n);
            if (x == n - 1) witness = false;
        }
        if (witness) return 0;
    }
    return 1;
}

uint64_t mps_multiplicative_order(uint64_t base, uint64_t prime) {
    uint64_t order = prime - 1;
    std::vector<uint64_t> f;
    distinct_factors(prime - 1, f);
    for (uint64_t q : f)
        while (order % q == 0 && powmod(base % prime, order / q, prime) == 1) order /= q;
    return order;
}

// run_bsgs (Math.hpp:377-486): n = n_start + cm*k, B = b^cm, B^k == -c * b^-n_start.
int mps_run_bsgs(uint64_t prime, const mps_work_item* it, mps_relation* out) {
    if (prime < 1000 || prime >= (1ull << 62) || (prime & 1) == 0) return bsgs_plain(prime, it, out);
    const uint64_t bm = mod_signed(it->b, prime);
    const uint64_t c0 = mod_signed(it->c, prime);
    const uint64_t t = c0 == 0 ? 0 : prime - c0;
    if (bm == 0) return bsgs_plain(prime, it, out);
    if (t == 0) return 0;
    const uint64_t N0 = it->N0 < 0 ? 0 : (uint64_t)it->N0;
    const uint64_t N1 = (uint64_t)it->N1;
    const uint64_t cm = it->cls_mod == 0 ? 1 : it->cls_mod;
    const uint64_t n_start = N0 + ((it->cls_res % cm) + cm - (N0 % cm)) % cm;
    if (n_start >= N1) return 0;
    uint64_t kcount = (N1 - n_start + cm - 1) / cm;
    if (kcount > prime) kcount = prime;

    const Mont mt(prime);
    const uint64_t b_m = mt.to(bm);
    uint64_t B = mt.pow(b_m, cm);
    uint64_t T = mt.to(t);
    if (n_start != 0) T = mt.mul(T, mt.pow(mt.pow(b_m, n_start), prime - 2));
    uint64_t stride = 1, parity = 0;
    const uint64_t g = gcd(cm, prime - 1);
    if (g >= 2) {
        if (mt.pow(T, (prime - 1) / g) != mt.one) return 0;
    } else {
        const int lb = mt.legendre(B), lt = mt.legendre(T);
        if (lb == 1) {
            if (lt == -1) return 0;
        } else {
            stride = 2;
            if (lt == -1) {
                parity = 1;
                T = mt.mul(T, mt.pow(B, prime - 2));
            }
            B = mt.mul(B, B);
        }
    }
    if (kcount <= parity) return 0;
    const uint64_t jcount = (kcount - parity + stride - 1) / stride;
    uint64_t m = isqrt_ceil(jcount);
    if (m == 0) m = 1;
    Table& tab = tls_table;
    tab.reset(m);
    uint64_t pw = mt.one, cycle = 0;
    for (uint64_t i = 0; i < m; ++i) {
        if (i > 0 && pw == mt.one) {
            cycle = i;
            break;
        }
        tab.insert(pw, (uint32_t)i);
        pw = mt.mul(pw, B);
    }
    uint64_t j = 0;
    bool found = false;
    uint32_t idx = 0;
    if (cycle) {
        found = tab.find(T, idx);
        j = idx;
    } else {
        const uint64_t step_inv = mt.pow(pw, prime - 2);
        uint64_t giant = T;
        for (uint64_t gs = 0; gs * m < jcount; ++gs) {
            if (tab.find(giant, idx)) {
                j = idx + gs * m;
                found = true;
                break;
            }
            giant = mt.mul(giant, step_inv);
        }
    }
    if (!found || j >= jcount) return 0;
    out->p = prime;
    out->M = n_start + (j * stride + parity) * cm;
    out->q = mps_multiplicative_order(bm, prime);
    return 1;
}

}  // extern "C"
