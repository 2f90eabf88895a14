// Host carry-less products of hgcd.cpp (512-bit output-block or 128-bit
// column schoolbook below the Karatsuba base, Karatsuba above it, block splitting
// for unbalanced operands, parallel half products from 1024 words)
// against a bit-serial shift-and-xor product.  Device products are off (gcd_dev_bits 0), so
// the device entry points below are never reached; they abort if called.
#include "host_poly.hpp"
#include "gf2t.h"
#include <cstdio>
#include <cstdlib>
#include <random>

using namespace gf2t_host;

extern "C" int gf2t_mul_host(const uint64_t *, uint64_t, const uint64_t *, uint64_t, uint64_t *) {
    abort();
}
extern "C" int gf2t_mat2_host(const uint64_t *const *, const uint64_t *, const uint64_t *, uint64_t,
                              const uint64_t *, uint64_t, uint64_t *, uint64_t *, size_t) {
    abort();
}
extern "C" int gf2t_mat2n_host(const uint64_t *const *, const uint64_t *, int,
                               const uint64_t *const *, const uint64_t *, uint64_t *const *,
                               size_t) {
    abort();
}
extern "C" size_t gf2t_mul_out_words(uint64_t, uint64_t) { abort(); }
extern "C" const char *gf2t_last_error(void) { return "no device in this check"; }
namespace gf2t_host {
bool square_fuse_enabled() { return true; }
}

static Poly bitserial(const Poly &a, const Poly &b) {
    Poly out(a.size() + b.size() + 1, 0);
    for (size_t jw = 0; jw < b.size(); jw++)
        for (int jb = 0; jb < 64; jb++) {
            if (!((b[jw] >> jb) & 1)) continue;
            for (size_t i = 0; i < a.size(); i++) {
                out[i + jw] ^= a[i] << jb;
                if (jb) out[i + jw + 1] ^= a[i] >> (64 - jb);
            }
        }
    trim(out);
    return out;
}

int main() {
    set_device_product_bits(0);
    std::mt19937_64 rng(7);
    auto rnd = [&](size_t n, int zero_every) {
        Poly p(n);
        for (size_t i = 0; i < n; i++) p[i] = (zero_every && i % zero_every == 1) ? 0 : rng();
        if (n) p[n - 1] |= 1ULL << (rng() % 64);   // top word nonzero
        return p;
    };
    const size_t sizes[][2] = {
        {1, 1},     {1, 2},    {2, 1},     {2, 2},     {3, 5},     {7, 7},    {23, 24},
        {24, 24},   {25, 25},  {31, 33},   {32, 32},   {33, 33},   {47, 49},  {48, 48},
        {49, 50},   {64, 65},  {1, 300},   {300, 1},   {5, 200},   {23, 1000}, {31, 1000},
        {32, 1000}, {33, 1000}, {96, 97},  {130, 129}, {333, 77},  {512, 512}, {700, 900},
        {1100, 1100}, {8, 8},   {9, 7},    {15, 17},   {16, 16},   {17, 9},   {40, 8},
        {8, 40},    {1025, 1025}, {2048, 2047}};
    int bad = 0, n = 0;
    for (auto &s : sizes)
        for (int zero_every : {0, 3}) {
            Poly a = rnd(s[0], zero_every), b = rnd(s[1], zero_every);
            Poly got = mul_fast(a, b);
            trim(got);
            if (!equal(got, bitserial(a, b))) {
                printf("mismatch %zu x %zu (zeros every %d)\n", s[0], s[1], zero_every);
                bad++;
            }
            n++;
        }
    // all-ones words: every partial product has its top bit set
    for (size_t w : {1, 24, 25, 64}) {
        Poly a(w, ~0ULL), b(w + 3, ~0ULL);
        Poly got = mul_fast(a, b);
        trim(got);
        if (!equal(got, bitserial(a, b))) {
            printf("mismatch all-ones %zu x %zu\n", w, w + 3);
            bad++;
        }
        n++;
    }
    if (bad) return 1;
    printf("host products ok (%d cases)\n", n);
    return 0;
}
