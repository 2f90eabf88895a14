// host_poly.cpp -- host GF(2)[x] arithmetic (see host_poly.hpp).
#include "host_poly.hpp"
#include <wmmintrin.h>
#include <smmintrin.h>
#include <algorithm>
#include <cstring>

namespace gf2t_host {

static inline void clmul(u64 a, u64 b, u64 &lo, u64 &hi) {
    __m128i p = _mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a),
                                     _mm_cvtsi64_si128((long long)b), 0);
    lo = (u64)_mm_cvtsi128_si64(p);
    hi = (u64)_mm_extract_epi64(p, 1);
}

int64_t deg_w(const u64 *a, size_t n) {
    for (size_t i = n; i-- > 0;)
        if (a[i]) return (int64_t)(i * 64 + 63 - __builtin_clzll(a[i]));
    return -1;
}
int64_t deg(const Poly &a) { return deg_w(a.data(), a.size()); }

void trim(Poly &a) {
    while (!a.empty() && a.back() == 0) a.pop_back();
}
bool is_zero(const Poly &a) { return deg(a) < 0; }

bool less(const Poly &a, const Poly &b) {
    size_t n = std::max(a.size(), b.size());
    for (size_t i = n; i-- > 0;) {
        u64 x = i < a.size() ? a[i] : 0, y = i < b.size() ? b[i] : 0;
        if (x != y) return x < y;
    }
    return false;
}
bool equal(const Poly &a, const Poly &b) { return !less(a, b) && !less(b, a); }

// a ^= b << sh over the words of b up to its degree
static inline void xor_shift_raw(u64 *a, const u64 *b, size_t bw, u64 sh) {
    size_t ws = (size_t)(sh >> 6);
    unsigned bs = (unsigned)(sh & 63);
    if (!bs) {
        for (size_t j = 0; j < bw; j++) a[j + ws] ^= b[j];
        return;
    }
    u64 carry = 0;
    for (size_t j = 0; j < bw; j++) {
        u64 v = b[j];
        a[j + ws] ^= (v << bs) | carry;
        carry = v >> (64 - bs);
    }
    a[bw + ws] ^= carry;
}

void xor_shift(Poly &a, const Poly &b, u64 sh) {
    int64_t db = deg(b);
    if (db < 0) return;
    size_t bw = (size_t)(db >> 6) + 1;
    size_t need = bw + (size_t)(sh >> 6) + 1;
    if (a.size() < need) a.resize(need, 0);
    xor_shift_raw(a.data(), b.data(), bw, sh);
}

Poly mul(const Poly &a, const Poly &b) {
    int64_t da = deg(a), db = deg(b);
    if (da < 0 || db < 0) return Poly();
    size_t aw = (size_t)(da >> 6) + 1, bw = (size_t)(db >> 6) + 1;
    Poly out(aw + bw + 1, 0);
    for (size_t i = 0; i < aw; i++) {
        if (!a[i]) continue;
        for (size_t j = 0; j < bw; j++) {
            u64 lo, hi;
            clmul(a[i], b[j], lo, hi);
            out[i + j] ^= lo;
            out[i + j + 1] ^= hi;
        }
    }
    trim(out);
    return out;
}

static inline u64 spread32(uint32_t v) {
    u64 x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFULL;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFULL;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0FULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

Poly sqr(const Poly &a) {
    int64_t da = deg(a);
    if (da < 0) return Poly();
    size_t aw = (size_t)(da >> 6) + 1;
    Poly out(2 * aw, 0);
    for (size_t i = 0; i < aw; i++) {
        out[2 * i] = spread32((uint32_t)a[i]);
        out[2 * i + 1] = spread32((uint32_t)(a[i] >> 32));
    }
    trim(out);
    return out;
}

// a <- a mod g by subtracting shifted copies of g from the top
void mod_inplace(Poly &a, const Poly &g) {
    int64_t dg = deg(g);
    int64_t da = deg(a);
    if (dg < 0 || da < dg) { trim(a); return; }
    size_t gw = (size_t)(dg >> 6) + 1;
    if (a.size() < (size_t)(da >> 6) + 2) a.resize((size_t)(da >> 6) + 2, 0);
    u64 *A = a.data();
    while (da >= dg) {
        xor_shift_raw(A, g.data(), gw, (u64)(da - dg));
        size_t top = (size_t)(da >> 6);
        while (true) {
            if (A[top]) { da = (int64_t)(top * 64 + 63 - __builtin_clzll(A[top])); break; }
            if (top == 0) { da = -1; break; }
            top--;
        }
    }
    a.resize((size_t)((dg >> 6) + 1));
    trim(a);
}

Poly divexact(const Poly &a, const Poly &g) {
    Poly r = a;
    int64_t dg = deg(g), dr = deg(r);
    if (dg < 0) return Poly();
    Poly q(dr >= dg ? (size_t)((dr - dg) >> 6) + 1 : 1, 0);
    size_t gw = (size_t)(dg >> 6) + 1;
    r.resize(r.size() + 2, 0);
    while (dr >= dg) {
        u64 sh = (u64)(dr - dg);
        q[sh >> 6] |= 1ULL << (sh & 63);
        xor_shift_raw(r.data(), g.data(), gw, sh);
        dr = deg_w(r.data(), (size_t)(dr >> 6) + 1);
    }
    trim(q);
    return q;
}

Poly gcd(const Poly &a0, const Poly &b0) {
    Poly a = a0, b = b0;
    trim(a);
    trim(b);
    if (deg(a) < deg(b)) std::swap(a, b);
    while (deg(b) >= 0) {
        mod_inplace(a, b);
        std::swap(a, b);
    }
    trim(a);
    return a;
}

Poly mulmod(const Poly &a, const Poly &b, const Poly &g) {
    Poly p = mul(a, b);
    mod_inplace(p, g);
    return p;
}

Poly sqrmod(const Poly &a, const Poly &g) {
    Poly p = sqr(a);
    mod_inplace(p, g);
    return p;
}

std::string hex(const Poly &a) {
    int64_t d = deg(a);
    if (d < 0) return "0";
    static const char *digits = "0123456789abcdef";
    std::string s;
    s.reserve((size_t)(d / 4) + 1);
    for (int64_t n = d / 4; n >= 0; n--)
        s.push_back(digits[(a[(size_t)(n >> 4)] >> ((n & 15) * 4)) & 0xF]);
    return s;
}

Poly from_words(const u64 *w, size_t n) {
    Poly p(w, w + n);
    trim(p);
    return p;
}

Poly trinomial(u64 r, u64 s) {
    Poly t((size_t)(r >> 6) + 1, 0);
    t[0] |= 1;
    t[s >> 6] |= 1ULL << (s & 63);
    t[r >> 6] |= 1ULL << (r & 63);
    return t;
}

static Poly x_plus(const Poly &h) {
    Poly t = h;
    if (t.empty()) t.push_back(0);
    t[0] ^= 2;
    trim(t);
    return t;
}

Modulus::Modulus(const Poly &g) : g_(g) {
    trim(g_);
    n_ = deg(g_);
    barrett_ = n_ >= 256;
    if (barrett_) {
        Poly x2n((size_t)((2 * n_) >> 6) + 1, 0);
        x2n[(size_t)((2 * n_) >> 6)] = 1ULL << ((2 * n_) & 63);
        // mu = x^(2n) div g by long division
        Poly r = x2n;
        int64_t dr = 2 * n_;
        mu_.assign((size_t)(n_ >> 6) + 1, 0);
        size_t gw = (size_t)(n_ >> 6) + 1;
        r.resize(r.size() + 2, 0);
        while (dr >= n_) {
            u64 sh = (u64)(dr - n_);
            mu_[sh >> 6] |= 1ULL << (sh & 63);
            xor_shift_words(r.data(), g_.data(), gw, sh);
            dr = deg_w(r.data(), (size_t)(dr >> 6) + 1);
        }
        trim(mu_);
    }
}

static Poly shr_bits(const Poly &a, u64 k) {
    size_t ws = (size_t)(k >> 6);
    unsigned bs = (unsigned)(k & 63);
    if (ws >= a.size()) return Poly();
    Poly out(a.size() - ws);
    for (size_t i = 0; i < out.size(); i++) {
        u64 lo = a[ws + i] >> bs;
        u64 hi = (bs && ws + i + 1 < a.size()) ? a[ws + i + 1] << (64 - bs) : 0;
        out[i] = lo | hi;
    }
    trim(out);
    return out;
}

Poly Modulus::reduce(const Poly &a0) const {
    Poly a = a0;
    trim(a);
    int64_t da = deg(a);
    if (da < n_) return a;
    if (!barrett_ || da >= 2 * n_) {
        mod_inplace(a, g_);
        return a;
    }
    // q = ((a div x^n) * mu) div x^n; a - q g has degree < n
    Poly q = shr_bits(mul_fast(shr_bits(a, (u64)n_), mu_), (u64)n_);
    Poly qg = mul_fast(q, g_);
    if (a.size() < qg.size()) a.resize(qg.size(), 0);
    for (size_t i = 0; i < qg.size(); i++) a[i] ^= qg[i];
    trim(a);
    if (deg(a) >= n_) mod_inplace(a, g_);
    return a;
}

Poly Modulus::sqr(const Poly &a) const { return reduce(gf2t_host::sqr(a)); }
Poly Modulus::mul(const Poly &a, const Poly &b) const { return reduce(mul_fast(a, b)); }

// The first k in [ka, kb] with gcd(x^(2^k) + x, g) != 1 is the least
// factor degree.  Blocks of 32 degrees are tested with one gcd of the
// product of their (x^(2^k) + x) mod g; only a block that hits is
// rescanned one degree at a time.
bool resolve_least(const Poly &g, u64 ka, u64 kb, const Poly &h0, u64 &d, Poly &mask) {
    Modulus G(g);
    Poly h = G.reduce(h0);
    const u64 B = 32;
    // With two or more irreducible factors (counted with multiplicity) the
    // least has degree at most deg g / 2; with one, that factor is g.  So
    // the walk stops at deg g / 2, and if it finds nothing g is irreducible
    // (its degree then lies in [ka, kb] like every factor's).
    const u64 gd = (u64)deg(g);
    const u64 kmax = std::min(kb, gd / 2);
    for (u64 k0 = ka; k0 <= kmax; k0 += B) {
        const u64 k1 = std::min(kmax, k0 + B - 1);
        const Poly start = h;
        Poly P{1};
        for (u64 k = k0; k <= k1; k++) {
            h = G.sqr(h);
            P = G.mul(P, x_plus(h));
        }
        if (deg(gcd_fast(P, g)) <= 0) continue;
        h = start;
        for (u64 k = k0; k <= k1; k++) {
            h = G.sqr(h);
            Poly t = gcd_fast(x_plus(h), g);
            if (deg(t) > 0) {
                d = k;
                mask = ((u64)deg(t) == k) ? t : edf_least(t, k);
                return !mask.empty();
            }
        }
        return false;   // the block's product hit but no single degree did
    }
    if (gd >= ka && gd <= kb) {   // no factor up to deg g / 2: g is irreducible
        d = gd;
        mask = g;
        return true;
    }
    return false;
}

// deterministic pseudo-random element of degree < n
static Poly pr_elem(u64 &state, int64_t n) {
    size_t w = (size_t)((n + 63) >> 6);
    Poly a(w, 0);
    for (size_t i = 0; i < w; i++) {
        state += 0x9E3779B97F4A7C15ULL;            // SplitMix64
        u64 z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        a[i] = z ^ (z >> 31);
    }
    if (n & 63) a[w - 1] &= (1ULL << (n & 63)) - 1;
    trim(a);
    return a;
}

// Equal-degree splitting over GF(2): for a in F[x]/t, the trace
// a + a^2 + ... + a^(2^(d-1)) is 0 or 1 modulo each degree-d factor, so
// gcd(trace, t) separates the factors on which it vanishes.  The set of
// factors is unique, so the numerically least one does not depend on
// the elements tried.
Poly edf_least(const Poly &t0, u64 d) {
    Poly t = t0;
    trim(t);
    int64_t dt = deg(t);
    if (dt <= 0) return Poly();
    if ((u64)dt == d) return t;
    u64 state = 0x243F6A8885A308D3ULL ^ (u64)dt;
    Modulus T(t);
    for (int tries = 0; tries < 200; tries++) {
        Poly a = pr_elem(state, dt);
        if (deg(a) < 1) continue;
        Poly tr = a, p = a;
        for (u64 i = 1; i < d; i++) {
            p = T.sqr(p);
            Poly acc = tr;
            xor_shift(acc, p, 0);
            trim(acc);
            tr.swap(acc);
        }
        Poly u = gcd_fast(tr, t);
        int64_t du = deg(u);
        if (du > 0 && du < dt) {
            Poly v = divexact(t, u);
            Poly f1 = edf_least(u, d), f2 = edf_least(v, d);
            if (f1.empty()) return f2;
            if (f2.empty()) return f1;
            return less(f1, f2) ? f1 : f2;
        }
    }
    return Poly();
}

u64 swan_rhigh(u64 r, u64 s, bool *swan_out) {
    bool swan = (((r & 7) == 1) || ((r & 7) == 7)) &&
                ((((s & 1) == 0) && (((2 * r) % s) != 0)) ||
                 (((s & 1) == 1) && (((2 * r) % (r - s)) != 0)));
    if (swan_out) *swan_out = swan;
    return swan ? (r / 3) : (r / 2);
}

}  // namespace gf2t_host
