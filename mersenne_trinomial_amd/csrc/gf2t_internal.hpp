// gf2t_internal.hpp -- shared device arithmetic and library plumbing for
// the gfx950 GF(2) trinomial kernels.  Everything here is carry-less
// integer work on 64-wide waves: XOR, shift, and 32x32 integer
// multiplies for the carry-less products (CDNA has no clmul instruction).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

typedef uint32_t u32;
typedef uint64_t u64;

#define GF2T_DEV __device__ __forceinline__
#define GF2T_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------- errors
namespace gf2t {
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
}  // namespace gf2t

#define GF2T_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e__ = (call);                                              \
        if (e__ != hipSuccess)                                                \
            return gf2t::hip_fail(e__, #call, __FILE__, __LINE__);            \
    } while (0)

#define GF2T_CHECK_LAUNCH(what)                                               \
    do {                                                                      \
        hipError_t e__ = hipGetLastError();                                   \
        if (e__ != hipSuccess)                                                \
            return gf2t::hip_fail(e__, what, __FILE__, __LINE__);             \
    } while (0)

// ------------------------------------------------------------ bit helpers
GF2T_HD u64 spread32(u32 v) {
    // interleave zeros: bit k of v -> bit 2k (squaring in GF(2)[x])
    u64 x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFULL;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFULL;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0FULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

// 64-bit window [q, q+64) of a bit array of nw words (zeros beyond)
GF2T_DEV u64 win64(const u64 *__restrict__ a, size_t nw, u64 q) {
    size_t w = (size_t)(q >> 6);
    unsigned sh = (unsigned)(q & 63);
    u64 lo = (w < nw) ? a[w] : 0;
    if (!sh) return lo;
    u64 hi = (w + 1 < nw) ? a[w + 1] : 0;
    return (lo >> sh) | (hi << (64 - sh));
}

// mask of bit positions [lc, hc) within a 64-bit word, 0 <= lc, hc <= 64
GF2T_HD u64 range_mask64(int lc, int hc) {
    if (hc <= lc) return 0;
    u64 hi = (hc >= 64) ? ~0ULL : ((1ULL << hc) - 1);
    u64 lo = (lc <= 0) ? 0 : ((1ULL << lc) - 1);
    return hi & ~lo;
}

// ---------------------------------------------- carry-less 32x32 -> 64
// Operands split into four bit classes (bits 4 apart); every partial
// integer product a_i*b_j has at most 8 set bits per column of its
// result class, so columns never carry into the next used position and
// the class bits are exact carry-less sums.
GF2T_HD u64 clmul32(u32 a, u32 b) {
    const u32 m0 = 0x11111111u, m1 = 0x22222222u, m2 = 0x44444444u,
              m3 = 0x88888888u;
    u64 a0 = a & m0, a1 = a & m1, a2 = a & m2, a3 = a & m3;
    u64 b0 = b & m0, b1 = b & m1, b2 = b & m2, b3 = b & m3;
    u64 c0 = (a0 * b0) ^ (a1 * b3) ^ (a2 * b2) ^ (a3 * b1);
    u64 c1 = (a0 * b1) ^ (a1 * b0) ^ (a2 * b3) ^ (a3 * b2);
    u64 c2 = (a0 * b2) ^ (a1 * b1) ^ (a2 * b0) ^ (a3 * b3);
    u64 c3 = (a0 * b3) ^ (a1 * b2) ^ (a2 * b1) ^ (a3 * b0);
    return (c0 & 0x1111111111111111ULL) ^ (c1 & 0x2222222222222222ULL) ^
           (c2 & 0x4444444444444444ULL) ^ (c3 & 0x8888888888888888ULL);
}

// GF(2^64) = GF(2)[x]/(x^64 + x^4 + x^3 + x + 1) multiply: Karatsuba
// over clmul32, then fold the high word twice (x^64 = x^4+x^3+x+1).
GF2T_HD u64 gf64_mul(u64 a, u64 b) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 lo = clmul32(a0, b0);
    u64 hi = clmul32(a1, b1);
    u64 mid = clmul32(a0 ^ a1, b0 ^ b1) ^ lo ^ hi;
    lo ^= mid << 32;
    hi ^= mid >> 32;
    u64 r = lo ^ hi ^ (hi << 1) ^ (hi << 3) ^ (hi << 4);
    u64 s = (hi >> 63) ^ (hi >> 61) ^ (hi >> 60);
    return r ^ s ^ (s << 1) ^ (s << 3) ^ (s << 4);
}

// ------------------------------------------------------- GF(2^d), d <= 32
// remainder of a carry-less product t (degree <= 2d-2) modulo q
GF2T_HD u32 fred(u64 t, u64 q, int d) {
    for (int b = 2 * d - 2; b >= d; b--) t ^= (q << (b - d)) & (0ULL - ((t >> b) & 1));
    return (u32)t;
}
GF2T_HD u32 fmul(u32 a, u32 b, u64 q, int d) { return fred(clmul32(a, b), q, d); }
GF2T_HD u32 fsqr(u32 a, u64 q, int d) { return fred(spread32(a), q, d); }

static inline size_t words_of(u64 bits) { return (size_t)((bits + 63) >> 6); }
