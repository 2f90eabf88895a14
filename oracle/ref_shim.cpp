// ref_shim.cpp -- extern "C" entry points over the reference's own host
// engine headers (src/gf2_cantor_engine.h, src/gf2_gcd.h), compiled where
// they lie in the reference tree by oracle/Makefile into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement against
// the reference's code; never linked into the product.
#include "gf2_cantor_engine.h"
#include "gf2_gcd.h"
#include <cstring>

extern "C" {

int ref_cantor_basis(uint64_t *out) {
    CantorBasis B;
    B.build();
    if (!B.verify()) return -1;
    memcpy(out, B.b, sizeof(B.b));
    return B.len;
}

uint64_t ref_gf64_mul(uint64_t a, uint64_t b) { return gf64_mul_portable(a, b); }

void ref_fft_fwd(uint64_t *f, int m) {
    CantorBasis B; B.build();
    CantorFFT F; F.init(m, B);
    F.fwd(f);
}

void ref_fft_inv(uint64_t *f, int m) {
    CantorBasis B; B.build();
    CantorFFT F; F.init(m, B);
    F.inv(f);
}

void ref_fft_mul(const uint64_t *a, uint64_t abits, const uint64_t *b,
                 uint64_t bbits, uint64_t *out) {
    CantorBasis B; B.build();
    gf2x_fft_mul(a, abits, b, bbits, out, B);
}

int64_t ref_gcd_naive(const uint64_t *a, size_t aw, const uint64_t *b,
                      size_t bw, uint64_t *g, size_t gcap) {
    std::vector<u64> G;
    u64 d = poly_gcd_naive(a, aw, b, bw, &G);
    if (g && gcap) {
        memset(g, 0, gcap * 8);
        memcpy(g, G.data(), (G.size() < gcap ? G.size() : gcap) * 8);
    }
    return (int64_t)d;
}

// poly_hex of the reference (src/gf2_gcd.h:96-105) into buf
int ref_poly_hex(const uint64_t *w, size_t nw, char *buf, size_t cap) {
    std::vector<u64> v(w, w + nw);
    std::string s = poly_hex(v);
    if (s.size() + 1 > cap) return -1;
    memcpy(buf, s.c_str(), s.size() + 1);
    return (int)s.size();
}

}  // extern "C"
