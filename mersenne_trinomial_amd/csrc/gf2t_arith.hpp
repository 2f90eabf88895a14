// gf2t_arith.hpp -- the carry-less arithmetic the gfx950 kernels run,
// written once for device and host: the kernels include it through
// gf2t_internal.hpp, and a host build of the same functions is checked
// against the oracle on the CPU (tests/test_kernel_logic.py).
#pragma once
#include <stdint.h>
#include <stddef.h>

#if defined(__HIP__) || defined(__HIPCC__)
#define GF2T_HD __host__ __device__ __forceinline__
#define GF2T_HDM __host__ __device__ __forceinline__
#else
#define GF2T_HD static inline
#define GF2T_HDM inline   // member functions
#endif
// global-memory pointers in device code (direct global loads with a
// scalar base and a 32-bit offset instead of flat loads)
#if defined(__HIP_DEVICE_COMPILE__)
#define GF2T_GLOBAL __attribute__((address_space(1)))
#else
#define GF2T_GLOBAL
#endif

#ifndef GF2T_U_TYPES
#define GF2T_U_TYPES
typedef uint32_t u32;
typedef uint64_t u64;
#endif

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
GF2T_HD u64 win64(const u64 *a, size_t nw, u64 q) {
    size_t w = (size_t)(q >> 6);
    unsigned sh = (unsigned)(q & 63);
    u64 lo = (w < nw) ? a[w] : 0;
    if (!sh) return lo;
    u64 hi = (w + 1 < nw) ? a[w + 1] : 0;
    return (lo >> sh) | (hi << (64 - sh));
}

// mask of bit positions [lc, hc) within a 64-bit word; lc and hc may lie
// anywhere (clamped to [0, 64]), and no shift count ever reaches 64
GF2T_HD u64 range_mask64(int lc, int hc) {
    if (lc < 0) lc = 0;
    if (hc > 64) hc = 64;
    if (hc <= lc) return 0;
    u64 hi = (hc == 64) ? ~0ULL : ((1ULL << hc) - 1);
    u64 lo = (lc == 0) ? 0 : ((1ULL << lc) - 1);
    return hi & ~lo;
}

// ---------------------------------------------- carry-less 32x32 -> 64
// Three-input XOR and bitwise select (m ? x : y per bit): one gfx950
// instruction each on a 32-bit half (v_bitop3_b32, truth tables 0x96 and
// 0xCA over inputs 0xF0, 0xCC, 0xAA); the same functions in plain C on
// the host.
GF2T_HD u32 xor3_32(u32 a, u32 b, u32 c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}
GF2T_HD u32 sel_32(u32 m, u32 x, u32 y) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(m, x, y, 0xCA);
#else
    return (m & x) | (~m & y);
#endif
}
GF2T_HD u64 xor3(u64 a, u64 b, u64 c) {
    return ((u64)xor3_32((u32)(a >> 32), (u32)(b >> 32), (u32)(c >> 32)) << 32) |
           xor3_32((u32)a, (u32)b, (u32)c);
}
GF2T_HD u64 sel64(u64 m, u64 x, u64 y) {
    return ((u64)sel_32((u32)(m >> 32), (u32)(x >> 32), (u32)(y >> 32)) << 32) |
           sel_32((u32)m, (u32)x, (u32)y);
}

// Operands split into four bit classes (bits 4 apart): the integer
// product a_i*b_j of class i of a and class j of b has its bits in class
// i + j mod 4, with at most 8 terms per column, so a column's count fits
// below the next class position and its low bit is the carry-less sum.
// Two such products whose i + j differ by 4 have column counts that
// peak 4 positions apart and sum to at most 15 (checked by brute force on
// all-ones operands), so they are added in one multiply-add
// (v_mad_u64_u32); the other products of a class are XORed in.  The
// classes are then merged by bit selects.  40 VALU instructions: 8
// masks, 16 multiplies, 10 XORs, 6 selects.  (A three-class form over
// 16-bit halves with 24-bit multiplies, v_mul_u32_u24, was checked
// bit-exact and made the transform passes 6% slower on gfx950.)
GF2T_HD u64 clmul32(u32 a, u32 b) {
    const u32 m0 = 0x11111111u, m1 = 0x22222222u, m2 = 0x44444444u,
              m3 = 0x88888888u;
    u64 a0 = a & m0, a1 = a & m1, a2 = a & m2, a3 = a & m3;
    u64 b0 = b & m0, b1 = b & m1, b2 = b & m2, b3 = b & m3;
    u64 c0 = xor3(a0 * b0 + a1 * b3, a2 * b2, a3 * b1);     // i + j: 0 + 4, 4, 4
    u64 c1 = (a0 * b1 + a2 * b3) ^ (a1 * b0 + a3 * b2);     // 1 + 5, 1 + 5
    u64 c2 = xor3(a0 * b2 + a3 * b3, a1 * b1, a2 * b0);     // 2 + 6, 2, 2
    u64 c3 = xor3(a0 * b3, a1 * b2, a2 * b1) ^ (a3 * b0);   // 3, 3, 3, 3
    return sel64(0x1111111111111111ULL, c0,
                 sel64(0x2222222222222222ULL, c1, sel64(0x4444444444444444ULL, c2, c3)));
}

// GF(2^64) = GF(2)[x]/(x^64 + x^4 + x^3 + x + 1) multiply: Karatsuba
// over clmul32, then fold the high word twice (x^64 = x^4+x^3+x+1): the
// first fold's overflow s is at most 4 bits, so the second stays in the
// low word.
GF2T_HD u64 gf64_mul(u64 a, u64 b) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 lo = clmul32(a0, b0);
    u64 hi = clmul32(a1, b1);
    u64 mid = xor3(clmul32(a0 ^ a1, b0 ^ b1), lo, hi);
    lo ^= mid << 32;
    hi ^= mid >> 32;
    u64 r = xor3(xor3(lo, hi, hi << 1), hi << 3, hi << 4);
    u32 h = (u32)(hi >> 32);
    u32 sv = xor3_32(h >> 31, h >> 29, h >> 28);
    u32 rl = xor3_32((u32)r, xor3_32(sv, sv << 1, sv << 3), sv << 4);
    return ((r >> 32) << 32) | rl;
}

// ------------------------------------------------------- GF(2^d), d <= 32
// remainder of a carry-less product t (degree <= 2d-2) modulo q
GF2T_HD u32 fred(u64 t, u64 q, int d) {
    for (int b = 2 * d - 2; b >= d; b--) t ^= (q << (b - d)) & (0ULL - ((t >> b) & 1));
    return (u32)t;
}
GF2T_HD u32 fmul(u32 a, u32 b, u64 q, int d) { return fred(clmul32(a, b), q, d); }
GF2T_HD u32 fsqr(u32 a, u64 q, int d) { return fred(spread32(a), q, d); }


// ------------------------------------------------ GF(2^d), 32 < d <= 40
// 64 x 64 -> 128-bit carry-less product (Karatsuba over clmul32)
GF2T_HD void clmul64(u64 a, u64 b, u64 &lo, u64 &hi) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 l = clmul32(a0, b0), h = clmul32(a1, b1);
    u64 mid = xor3(clmul32(a0 ^ a1, b0 ^ b1), l, h);
    lo = l ^ (mid << 32);
    hi = h ^ (mid >> 32);
}
// remainder of the 128-bit product (degree <= 2d-2) modulo q (degree d)
GF2T_HD u64 fred64(u64 lo, u64 hi, u64 q, int d) {
    for (int b = 2 * d - 2; b >= d; b--) {
        u64 bit = b >= 64 ? (hi >> (b - 64)) & 1 : (lo >> b) & 1;
        if (!bit) continue;
        int sh = b - d;
        lo ^= q << sh;
        if (sh) hi ^= q >> (64 - sh);
    }
    return lo;
}
GF2T_HD u64 fmul64(u64 a, u64 b, u64 q, int d) {
    u64 lo, hi;
    clmul64(a, b, lo, hi);
    return fred64(lo, hi, q, d);
}
// strict minimum of its rotations within d <= 63 bits (Lyndon word)
GF2T_HD bool is_lyndon64(u64 j, int d, u64 M) {
    u64 jj = j;
    for (int i = 1; i < d; i++) {
        jj = ((jj << 1) | (jj >> (d - 1))) & M;
        if (jj <= j) return false;
    }
    return true;
}
GF2T_HD u64 gcd64(u64 a, u64 b) {
    while (b) { u64 t = a % b; a = b; b = t; }
    return a;
}
// inverse of a modulo m (gcd(a, m) = 1, 2 <= m < 2^62)
GF2T_HD u64 inv64(u64 a, u64 m) {
    u64 r0 = m, r1 = a % m;
    long long t0 = 0, t1 = 1;
    while (r1) {
        u64 q = r0 / r1;
        u64 r2 = r0 - q * r1;
        long long t2 = t0 - (long long)q * t1;
        r0 = r1; r1 = r2; t0 = t1; t1 = t2;
    }
    if (t0 < 0) t0 += (long long)m;
    return (u64)t0;
}
// a * b mod m for a, b < m < 2^62 (shift-and-add, no 128-bit division)
GF2T_HD u64 mulmod64(u64 a, u64 b, u64 m) {
    u64 r = 0;
    a %= m;
    while (b) {
        if (b & 1) { r += a; if (r >= m) r -= m; }
        a <<= 1;
        if (a >= m) a -= m;
        b >>= 1;
    }
    return r;
}
// Berlekamp-Massey minimal polynomial as minpoly_bm, for 32 < d <= 40:
// the trace sequence is generated on the fly (2d terms do not fit a word)
GF2T_HD u64 minpoly_bm64(u64 beta, u64 q, int d, u64 tmask) {
    u64 C = 1, B = 1, W = 0, pw = 1;
    int L = 0, m = 1;
    for (int n = 0; n < 2 * d; n++) {
        u64 a = (u64)(__builtin_popcountll(pw & tmask) & 1);
        pw = fmul64(pw, beta, q, d);
        W = (W << 1) | a;                               // bit j = a_{n-j}
        if (!(__builtin_popcountll(C & W) & 1)) { m++; continue; }
        if (m > 63) return 0;   // cannot happen at linear complexity <= 40
        u64 T = C;
        C ^= B << m;
        if (2 * L <= n) { L = n + 1 - L; B = T; m = 1; }
        else m++;
    }
    if (L != d) return 0;
    u64 mask = 0;
    for (int k = 0; k <= d; k++) mask |= ((C >> (d - k)) & 1) << k;
    return mask;
}

// ------------------------------------------- squaring + trinomial fold
// Byte permute: byte k of the result is byte (sel_k & 7) of the 8-byte
// value hi:lo (v_perm_b32 with selectors 0..7, the only ones used here).
GF2T_HD u32 perm8(u32 hi, u32 lo, u32 sel) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(hi, lo, sel);
#else
    const u64 v = ((u64)hi << 32) | lo;
    u32 out = 0;
    for (int k = 0; k < 4; k++)
        out |= (u32)((v >> (8 * ((sel >> (8 * k)) & 7))) & 0xFF) << (8 * k);
    return out;
#endif
}
// low 32 bits of (hi:lo) >> (sh & 31) (v_alignbit_b32)
GF2T_HD u32 alignbit(u32 hi, u32 lo, u32 sh) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, sh);
#else
    return (u32)((((u64)hi << 32) | lo) >> (sh & 31));
#endif
}
// (a >> sh) | (b << (64 - sh)) for 0 <= sh < 64 (sh = 0 gives a)
GF2T_HD u64 funnel64(u64 a, u64 b, u32 sh) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u32 lo, hi;
    if (sh < 32) { lo = alignbit(a1, a0, sh); hi = alignbit(b0, a1, sh); }
    else { lo = alignbit(b0, a1, sh); hi = alignbit(b1, b0, sh); }
    return ((u64)hi << 32) | lo;
}
// Spread of 32 bits into 64 (bit k -> bit 2k) by table lookup: the low
// three bits of each nibble select its spread byte from the eight-entry
// table {00 01 04 05 10 11 14 15} with one byte permute (four nibbles at
// a time), the nibble's top bit lands at bit 6 of that byte, and two
// more permutes interleave the bytes of the even and the odd nibbles.
GF2T_HD u64 spread32p(u32 x) {
    const u32 T_HI = 0x15141110u, T_LO = 0x05040100u;
    const u32 e = perm8(T_HI, T_LO, x & 0x07070707u) | ((x & 0x08080808u) << 3);
    const u32 o = perm8(T_HI, T_LO, (x >> 4) & 0x07070707u) | ((x >> 1) & 0x40404040u);
    const u32 lo = perm8(o, e, 0x05010400u);   // e0 o0 e1 o1
    const u32 hi = perm8(o, e, 0x07030602u);   // e2 o2 e3 o3
    return ((u64)hi << 32) | lo;
}
GF2T_HD void spread64(u64 v, u64 &lo, u64 &hi) {
    lo = spread32p((u32)v);
    hi = spread32p((u32)(v >> 32));
}
// Bit interleave: bit k of e -> bit 2k, bit k of o -> bit 2k + 1, i.e.
// spread(e) + x * spread(o).  The same nibble lookups as spread32p, with
// a second table (the spreads shifted left by one) for o, so each byte
// of the result is one lookup of e or'ed with one of o.
GF2T_HD u64 interleave32(u32 e, u32 o) {
    const u32 TE_HI = 0x15141110u, TE_LO = 0x05040100u;
    const u32 TO_HI = 0x2A282220u, TO_LO = 0x0A080200u;
    const u32 lo_n = perm8(TE_HI, TE_LO, e & 0x07070707u) | ((e & 0x08080808u) << 3) |
                     perm8(TO_HI, TO_LO, o & 0x07070707u) | ((o & 0x08080808u) << 4);
    const u32 hi_n = perm8(TE_HI, TE_LO, (e >> 4) & 0x07070707u) | ((e >> 1) & 0x40404040u) |
                     perm8(TO_HI, TO_LO, (o >> 4) & 0x07070707u) | (o & 0x80808080u);
    const u32 lo = perm8(hi_n, lo_n, 0x05010400u);
    const u32 hi = perm8(hi_n, lo_n, 0x07030602u);
    return ((u64)hi << 32) | lo;
}

// Word sources of the squaring.  src(win, k) is word k of the input,
// zero at and beyond nw; src.raw32(win, q) reads 32-bit word q unchecked,
// for the region fast path whose reads all lie inside the nw words (see
// below).
// `win` names the fold window a read serves (0 = SP[0,r), 1..4 = the
// four shifted windows): a plain array ignores it, the two-squaring
// kernel's tile keeps one LDS copy per window.
template <bool W12>
struct SqArrayT {
    const u64 *p;
    size_t nw;
    GF2T_HDM u64 operator()(int, size_t k) const { return k < nw ? p[k] : 0; }
    // 32-bit words q and q + 1 of the input, unchecked, and word q + 2
    // when sh != 0, the three as one 12-byte load when W12.  Without W12,
    // word q + 1 is read again when sh = 0 (the funnel shift does not use
    // it), so no read passes the window's last word.  W12 reads word
    // q + 2 also when sh = 0 (its value unused): inside the region fast
    // path a window's last bit is at most r - 1, so 32 (q + 2) <= r and
    // word q + 2 lies inside the 2 nw words unless r is a multiple of 64
    // (r = 64 (nw - 1) + t with t < 64 gives floor(r / 32) <= 2 nw - 1).
    // No branch on q's alignment or on sh: the loads of all the windows
    // of a pair stay in flight together.  32-bit byte offsets: residues
    // are below 2^29 words.
    GF2T_HDM const GF2T_GLOBAL u32 *at(u32 q) const {
        return (const GF2T_GLOBAL u32 *)((const GF2T_GLOBAL char *)p + (u32)(q << 2));
    }
    GF2T_HDM void raw3(int, u32 q, u32 sh, u32 &a, u32 &b, u32 &c) const {
        const GF2T_GLOBAL u32 *w = at(q);
        a = w[0];
        b = w[1];
        c = W12 ? w[2] : w[1 + (sh != 0)];
    }
};
using SqArray = SqArrayT<false>;
// whether the 12-byte window loads are in bounds for degree r
GF2T_HD bool sq_w12_ok(u64 r) { return (r & 63) != 0; }

// 128-bit window of the virtual spread SP of the source at SP position
// P, keeping only SP indices in [lo, hi).
template <class Src>
GF2T_HD void vspread128(const Src &src, int win, u64 P, u64 lo, u64 hi, u64 &w0, u64 &w1) {
    if (P >= hi || P + 128 <= lo) { w0 = 0; w1 = 0; return; }
    const u64 q = (P + 1) >> 1;
    const size_t w = (size_t)(q >> 6);
    const u32 sh = (u32)(q & 63);
    u64 v = src(win, w);
    if (sh) v = funnel64(v, src(win, w + 1), sh);
    u64 a, b;
    spread64(v, a, b);
    if (P & 1) { a <<= 1; b <<= 1; }
    int lc = lo > P ? (int)(lo - P) : 0;
    int hc = (hi - P) >= 128 ? 128 : (int)(hi - P);
    w0 = a & range_mask64(lc, hc);
    w1 = b & range_mask64(lc - 64, hc - 64);
}

// output words 2i, 2i+1 of in^2 mod x^r + x^s + 1:
//   SP[0,r) + SP[r,2r-1)>>r + SP[r,2r-s)>>(r-s)
//   + SP[2r-s,2r-1)>>(2r-s) + SP[2r-s,2r-1)>>(2r-2s)
template <class Src>
GF2T_HD void square_fold_pair_src(const Src &src, u64 r, u64 s, size_t i, u64 &o0, u64 &o1) {
    const u64 b = (u64)i << 7;
    u64 t0, t1;
    vspread128(src, 0, b, 0, r, o0, o1);
    vspread128(src, 1, b + r, r, 2 * r - 1, t0, t1);
    o0 ^= t0; o1 ^= t1;
    vspread128(src, 2, b + r - s, r, 2 * r - s, t0, t1);
    o0 ^= t0; o1 ^= t1;
    vspread128(src, 3, b + 2 * r - s, 2 * r - s, 2 * r - 1, t0, t1);
    o0 ^= t0; o1 ^= t1;
    vspread128(src, 4, b + 2 * r - 2 * s, 2 * r - s, 2 * r - 1, t0, t1);
    o0 ^= t0; o1 ^= t1;
}
GF2T_HD void square_fold_pair(const u64 *in, size_t nw, u64 r, u64 s, size_t i, u64 &o0,
                              u64 &o1) {
    square_fold_pair_src(SqArray{in, nw}, r, s, i, o0, o1);
}

// Window k of the squaring reads the 64 input bits starting at input bit
// 64i + Q_k (i = output pair index) and, when the spread position is odd,
// shifts the spread left by one.  Q_k, and so the word offset, the bit
// shift and the parity, are the same for every pair of one residue.
struct SqWin {
    u32 off32;    // 32-bit word offset of Q_k (r < 2^34, so it fits)
    u32 sh;       // Q_k mod 32
    GF2T_HDM u32 off() const { return off32 >> 1; }   // 64-bit word offset
};
struct SqConst {
    u64 r, s;
    size_t nw;
    SqWin w2, w3, w4, w5;
};
GF2T_HD SqWin sq_win(u64 P0) {   // P0 = window offset in spread positions (b = 0)
    u64 Q = (P0 + 1) >> 1;
    SqWin w;
    w.off32 = (u32)(Q >> 5);
    w.sh = (u32)(Q & 31);
    return w;
}
GF2T_HD SqConst sq_const(u64 r, u64 s) {
    SqConst c;
    c.r = r; c.s = s; c.nw = (size_t)((r + 63) >> 6);
    c.w2 = sq_win(r);
    c.w3 = sq_win(r - s);
    c.w4 = sq_win(2 * r - s);
    c.w5 = sq_win(2 * r - 2 * s);
    return c;
}
// Output pair i of in^2 mod T.  Away from the points s, 2s and r every
// window is either fully inside or fully outside its clip range:
//   [0, s)  : SP[0,r) + SP[r,..)>>r + SP[2r-s,..)>>(2r-s)          (W1 W2 W4)
//   [s, 2s) : W1 + W2 + SP[r,2r-s)>>(r-s) + SP[2r-s,..)>>(2r-2s)   (W1 W2 W3 W5)
//   [2s, r) : W1 + W2 + W3
// so those pairs skip all masking; pairs straddling s, 2s or r take the
// clipped general path.  In these three regions every window's last
// input bit is at most r - 1 (e.g. W2 in [2s, r): 64i + (r+1)/2 + 63
// with 128i + 128 <= r), so each read, the funnel shift's second word
// included, lies inside the nw input words.
GF2T_HD int sq_region(u64 b, u64 b1, u64 s, u64 r) {   // bit range [b, b1)
    if (b1 <= s) return 0;
    if (b >= s && b1 <= 2 * s) return 1;
    if (b >= 2 * s && b1 <= r) return 2;
    return 3;
}
// The 64 input bits from 32-bit word q on, shifted right by sh (< 32):
// three words and two funnel shifts (the third word is not read when
// sh = 0, where alignbit returns its low operand).  qodd = q & 1 is the
// same for every pair of a window.
template <class Src>
GF2T_HD u64 sq_winval(const Src &src, int win, u32 q, u32 qodd, u32 sh) {
    (void)qodd;
    u32 a, b, c;
    src.raw3(win, q, sh, a, b, c);
    return ((u64)alignbit(c, b, sh) << 32) | alignbit(b, a, sh);
}
// Output pair i in region REG (0..2).  Spreading is linear and a window
// at an odd spread position is the spread shifted by one, so the window
// values are summed by parity first (E: even windows, O: odd) and the
// pair is the interleave of E and O.  The parities follow from those of
// r and s (W2: r, W3: r - s, W4: s, W5: even), so they are template
// arguments and the sums cost one xor per window.
template <int REG, int RP, int SP, class Src>
GF2T_HD void sq_pair_t(const Src &src, const SqConst &C, u32 i, u64 &o0, u64 &o1) {
    u64 E = sq_winval(src, 0, 2 * i, 0, 0), O = 0;
    const u64 v2 = sq_winval(src, 1, 2 * i + C.w2.off32, C.w2.off32 & 1, C.w2.sh);
    if (RP) O ^= v2; else E ^= v2;
    if (REG == 0) {
        const u64 v4 = sq_winval(src, 3, 2 * i + C.w4.off32, C.w4.off32 & 1, C.w4.sh);
        if (SP) O ^= v4; else E ^= v4;
    } else {
        const u64 v3 = sq_winval(src, 2, 2 * i + C.w3.off32, C.w3.off32 & 1, C.w3.sh);
        if (RP ^ SP) O ^= v3; else E ^= v3;
        if (REG == 1) E ^= sq_winval(src, 4, 2 * i + C.w5.off32, C.w5.off32 & 1, C.w5.sh);
    }
    o0 = interleave32((u32)E, (u32)O);
    o1 = interleave32((u32)(E >> 32), (u32)(O >> 32));
}
// The twelve (region 0..2) x (r, s parity) variants: sq_variant calls
// f(SqV<REG, RP, SP>{}) for the one that applies, so a caller picks it
// once for a whole tile or window (the arguments are uniform there).
template <int REG, int RP, int SP>
struct SqV {
    template <class Src>
    GF2T_HDM void pair(const Src &src, const SqConst &C, u32 i, u64 &o0, u64 &o1) const {
        sq_pair_t<REG, RP, SP>(src, C, i, o0, o1);
    }
};
template <class F>
GF2T_HD void sq_variant(int region, const SqConst &C, F &&f) {
    switch (region * 4 + (int)((C.r & 1) << 1) + (int)(C.s & 1)) {
    case 0: f(SqV<0, 0, 0>{}); break;
    case 1: f(SqV<0, 0, 1>{}); break;
    case 2: f(SqV<0, 1, 0>{}); break;
    case 3: f(SqV<0, 1, 1>{}); break;
    case 4: f(SqV<1, 0, 0>{}); break;
    case 5: f(SqV<1, 0, 1>{}); break;
    case 6: f(SqV<1, 1, 0>{}); break;
    case 7: f(SqV<1, 1, 1>{}); break;
    case 8: f(SqV<2, 0, 0>{}); break;
    case 9: f(SqV<2, 0, 1>{}); break;
    case 10: f(SqV<2, 1, 0>{}); break;
    default: f(SqV<2, 1, 1>{}); break;
    }
}
// The same pair in two stages, for a caller that issues the loads of
// several pairs before combining any: the raw window values of pair i
// (region reg 0..2, uniform), then their sums by parity and interleave.
struct SqRawPair {
    u64 x, v2, vb, v5;   // W1, W2, W4 (region 0) or W3, W5 (region 1, else 0)
};
template <class Src>
GF2T_HD SqRawPair sq_raw_pair(const Src &src, const SqConst &C, u32 i, int reg) {
    SqRawPair w;
    w.x = sq_winval(src, 0, 2 * i, 0, 0);
    w.v2 = sq_winval(src, 1, 2 * i + C.w2.off32, C.w2.off32 & 1, C.w2.sh);
    const u32 boff = reg == 0 ? C.w4.off32 : C.w3.off32, bsh = reg == 0 ? C.w4.sh : C.w3.sh;
    w.vb = sq_winval(src, reg == 0 ? 3 : 2, 2 * i + boff, boff & 1, bsh);
    w.v5 = 0;
    if (reg == 1) w.v5 = sq_winval(src, 4, 2 * i + C.w5.off32, C.w5.off32 & 1, C.w5.sh);
    return w;
}
GF2T_HD void sq_combine_pair(const SqRawPair &w, const SqConst &C, int reg, u64 &o0, u64 &o1) {
    const u64 m2 = (C.r & 1) ? ~0ULL : 0;
    const u64 mb = ((reg == 0 ? C.s : C.r ^ C.s) & 1) ? ~0ULL : 0;
    const u64 O = (w.v2 & m2) ^ (w.vb & mb);
    const u64 E = w.x ^ w.v2 ^ w.vb ^ w.v5 ^ O;
    o0 = interleave32((u32)E, (u32)O);
    o1 = interleave32((u32)(E >> 32), (u32)(O >> 32));
}
// one pair in region 0..2
template <class Src>
GF2T_HD void sq_pair_region(const Src &src, const SqConst &C, u32 i, int region, u64 &o0,
                            u64 &o1) {
    sq_variant(region, C, [&](auto v) { v.pair(src, C, i, o0, o1); });
}
template <class Src>
GF2T_HD void square_fold_pair_fast_src(const Src &src, const SqConst &C, size_t i, u64 &o0,
                                       u64 &o1) {
    const u64 b = (u64)i << 7;
    const int region = sq_region(b, b + 128, C.s, C.r);
    if (region == 3) square_fold_pair_src(src, C.r, C.s, i, o0, o1);
    else sq_pair_region(src, C, (u32)i, region, o0, o1);
}
template <bool W12 = false>
GF2T_HD void square_fold_pair_fast(const u64 *in, const SqConst &C, size_t i, u64 &o0, u64 &o1) {
    square_fold_pair_fast_src(SqArrayT<W12>{in, C.nw}, C, i, o0, o1);
}
// region shared by the pairs [i0, i1) (3: mixed or clipped)
GF2T_HD int sq_tile_region(const SqConst &C, size_t i0, size_t i1) {
    return sq_region((u64)i0 << 7, (u64)i1 << 7, C.s, C.r);
}

// ---- two squarings per pass: F -> G = F^2 -> H = G^2 (mod T)
// A tile of output pairs [i0, i0 + SQ2_TP) of H reads G only through
// the windows of its pairs: window k covers G words [i0 + off_k,
// i0 + off_k + SQ2_TP] (the last one for the funnel shift).  With
// SQ2_TP <= 510 those words lie inside the 256 G pairs from
// (i0 + off_k) >> 1 on, so each of the 256 threads computes one G pair
// per window, straight from F, into that window's copy (512 words).
// Tiles inside one region need three or four of the five windows; a
// tile that straddles s, 2s or r takes all five.  The tile's H pairs are
// then formed from those copies: F is read once per two squarings.
// SQ2_TP = 504 = 63 * 8 leaves room for the H phase to map its threads
// from i0 rounded down to 8 pairs (one 128-byte line), so that each
// wave's 1 KiB of non-temporal stores covers whole lines: the block's
// 512 pair slots cover (i0 & 7) + width <= 7 + 504 pairs.
constexpr int SQ2_WINS = 5;
constexpr u32 SQ2_TP = 504;
constexpr u32 SQ2_GP = 256;          // G pairs per window
constexpr u32 SQ2_LEN = 2 * SQ2_GP;  // G words per window copy
GF2T_HD u32 sq2_off(const SqConst &C, int k) {   // word offset of window k
    return k == 0 ? 0 : k == 1 ? C.w2.off() : k == 2 ? C.w3.off() : k == 3 ? C.w4.off() : C.w5.off();
}
GF2T_HD u32 sq2_need(int region) {   // windows read by a tile of this region
    return region == 0 ? 0x0Bu : region == 1 ? 0x17u : region == 2 ? 0x07u : 0x1Fu;
}
// first G pair of window k for the tile at i0
GF2T_HD u64 sq2_p0(const SqConst &C, size_t i0, int k) { return ((u64)i0 + sq2_off(C, k)) >> 1; }
// G pair p, words 2p and 2p + 1 of G (zero at and beyond nw), by the
// clipped path: for a window whose G pairs straddle s, 2s or r
// (sq_tile_region(C, p0, p0 + SQ2_GP) == 3)
template <class Src>
GF2T_HD void sq2_g_pair_clipped(const Src &F, const SqConst &C, u64 p, u64 &o0, u64 &o1) {
    o0 = o1 = 0;
    if (2 * p < C.nw) square_fold_pair_fast_src(F, C, (size_t)p, o0, o1);
    if (2 * p + 1 >= C.nw) o1 = 0;
}
// the H phase reads G through the tile's window copies: window k's copy
// holds G words from base_k = 2 p0_k on, at word at_k of the LDS area
struct Sq2Tile {
    const u64 *win;
    u32 base[SQ2_WINS];
    u32 at[SQ2_WINS];
    GF2T_HDM u64 operator()(int k, size_t w) const {
        return win[at[k] + (u32)(w - base[k])];
    }
    // as SqArray::raw3, from the copy's 32-bit words q, q + 1 and q + 2
    // for every shift: the copy holds the word after a window's last one
    // (the funnel shift's word when sh != 0), so q + 2 is always inside it
    GF2T_HDM void raw3(int k, u32 q, u32 sh, u32 &a, u32 &b, u32 &c) const {
        (void)sh;
        const u32 *w = (const u32 *)win + (2 * at[k] + (u32)(q - 2 * base[k]));
        a = w[0];
        b = w[1];
        c = w[2];
    }
};
// what the G phase of one tile computes, and where each window's copy
// lives: compute = windows formed from F; every needed window not in
// compute is a copy the block already holds
struct Sq2Plan {
    u32 compute;
    u32 base[SQ2_WINS];
    u32 at[SQ2_WINS];
};
// a stand-alone tile: every needed window computed, window k in slot k
GF2T_HD Sq2Plan sq2_plan(const SqConst &C, size_t i0, u32 need) {
    Sq2Plan p;
    p.compute = need;
#pragma unroll
    for (int k = 0; k < SQ2_WINS; k++) {
        p.base[k] = (u32)(2 * sq2_p0(C, i0, k));
        p.at[k] = (u32)k * SQ2_LEN;
    }
    return p;
}
GF2T_HD Sq2Tile sq2_tile(const u64 *win, const Sq2Plan &p) {
    Sq2Tile t;
    t.win = win;
#pragma unroll
    for (int k = 0; k < SQ2_WINS; k++) {
        t.base[k] = p.base[k];
        t.at[k] = p.at[k];
    }
    return t;
}

// ---- chained tiles of the two-squaring pass
// A tile of output pairs [i0, i0 + width) reads window k's G words
// [i0 + off_k, i0 + off_k + width].  With delta = off2 - off3 (words),
// window 3 of the tile at i0 + delta reads exactly the words window 2 of
// the tile at i0 read (same base), and window 5 of the tile at i0 +
// delta reads window 4's words of the tile at i0 shifted by
// off5 - off4 + delta, a word or two at most.  So the output pairs are
// cut into work units: a unit is up to SQ2_CHAIN tiles i0, i0 + delta,
// i0 + 2 delta, ... of one column, and its block keeps the previous
// tile's window-2 and window-4 copies in LDS as the current tile's
// windows 3 and 5.  Chained tiles are at most SQ2_CW = SQ2_TP - 2 pairs
// wide, so one 512-word copy of window 4 can cover both its own words
// and the next tile's window 5 (the copy's base is placed for both;
// where it cannot cover, window 5 is formed).  Each period of delta
// pairs is split into ncols columns of wlo or wlo + 1 pairs.  In [2s, r) a tile
// then forms windows 1 and 2 only, in [s, 2s) likewise, where plain
// tiles form three and four.  For delta < SQ2_CW (s below about
// 128 SQ2_CW) units are the plain consecutive tiles of SQ2_TP pairs.
// LDS slots: 0 window 1; 1, 2 windows 2 (current) and 3 (the previous
// window 2), alternating from tile to tile; 3, 4 likewise windows 4, 5.
constexpr u32 SQ2_CHAIN = 8;
constexpr u32 SQ2_CW = SQ2_TP - 2;
// pair and word indices are below 2^28 (r < 2^34), so all of it is u32.
// The layout is computed once per residue on the host (sq2_chain, with
// its divisions); a block derives its tiles from it with one 32-bit
// division.
struct Sq2Chain {
    u32 np;       // output pairs
    u32 period;   // distance between consecutive tiles of a unit
    u32 ncols;    // columns per period
    u32 len;      // tiles per unit
    u32 chained;
    u32 wlo, wrem;   // column widths: wlo + 1 for the first wrem columns, wlo after
    u32 q, rm;       // np = q period + rm
};
GF2T_HD u32 sq2_col_start(const Sq2Chain &h, u32 c) {
    return c * h.wlo + (c < h.wrem ? c : h.wrem);
}
// tiles of column c: start + j period < np for j < q, and j = q too when
// the column starts below rm
GF2T_HD u32 sq2_col_tiles(const Sq2Chain &h, u32 c) {
    return h.q + (sq2_col_start(h, c) < h.rm ? 1u : 0u);
}
inline Sq2Chain sq2_chain(const SqConst &C) {   // host
    Sq2Chain h;
    h.np = (u32)((C.nw + 1) / 2);
    const u32 d = C.w2.off() - C.w3.off();
    h.chained = d >= SQ2_CW;
    h.period = h.chained ? d : SQ2_TP;
    h.ncols = h.chained ? (d + SQ2_CW - 1) / SQ2_CW : 1u;
    h.len = h.chained ? SQ2_CHAIN : 1u;
    h.wlo = h.period / h.ncols;
    h.wrem = h.period % h.ncols;
    h.q = h.np / h.period;
    h.rm = h.np % h.period;
    return h;
}
// units of one residue (column 0 has the most tiles)
inline u64 sq2_units(const Sq2Chain &h) {
    return (u64)h.ncols * ((sq2_col_tiles(h, 0) + h.len - 1) / h.len);
}
// the tiles [j0, j1) of unit u, in its column: tile j covers the pairs
// [i0, i1) with i0 = start + j period, i1 = min(i0 + width, np);
// false for a unit past the residue's last
struct Sq2Unit {
    u32 start, width, j0, j1;
};
GF2T_HD bool sq2_unit(const Sq2Chain &h, u32 u, Sq2Unit &w) {
    const u32 piece = h.ncols == 1 ? u : u / h.ncols;
    const u32 c = u - piece * h.ncols;
    const u32 nt = sq2_col_tiles(h, c);
    w.start = sq2_col_start(h, c);
    w.width = h.wlo + (c < h.wrem ? 1u : 0u);
    w.j0 = piece * h.len;
    w.j1 = w.j0 + h.len < nt ? w.j0 + h.len : nt;
    return w.j0 < nt;
}
// plan of the unit's tile number t (t = j - j0) at i0, `width` pairs of
// its column wide, needing the windows `need`; the previous tile of the
// unit needed prev_need and held its window 4 from word prev_base4 on
GF2T_HD Sq2Plan sq2_chain_plan(const SqConst &C, const Sq2Chain &h, u32 i0, u32 width,
                               u32 need, u32 t, u32 prev_need, u32 prev_base4) {
    Sq2Plan p;
    const u32 cur = t & 1, prv = cur ^ 1u;
#pragma unroll
    for (int k = 0; k < SQ2_WINS; k++) p.base[k] = (i0 + sq2_off(C, k)) & ~1u;
    p.at[0] = 0;
    p.at[1] = (1 + cur) * SQ2_LEN;
    p.at[2] = (1 + prv) * SQ2_LEN;
    p.at[3] = (3 + cur) * SQ2_LEN;
    p.at[4] = (3 + prv) * SQ2_LEN;
    p.compute = need;
    if (!h.chained) return p;
    const u32 off4 = sq2_off(C, 3), off5 = sq2_off(C, 4);
    if (need & 8u) {   // window 4: one copy for its words and the next tile's window 5
        const u32 own = i0 + off4, nxt = i0 + h.period + off5;
        const u32 lo = own < nxt ? own : nxt, hi = own < nxt ? nxt : own;
        if (hi + width < (lo & ~1u) + SQ2_LEN) p.base[3] = lo & ~1u;
    }
    if (t > 0) {
        // every tile forms window 2; window 3 has the same base
        if (need & 4u) p.compute &= ~4u;
        const u32 lo = i0 + off5;
        if ((need & 16u) && (prev_need & 8u) && lo >= prev_base4 &&
            lo + width < prev_base4 + SQ2_LEN) {
            p.compute &= ~16u;
            p.base[4] = prev_base4;
        }
    }
    return p;
}

// ---- XCD-aware visiting order of the output tiles of one residue
// Window 2 of output pair i and window 3 of pair i + delta read the same
// input words, and so do windows 4 and 5 (delta = off2 - off3 ~ s/128
// pairs).  Launches hand workgroups to the 8 XCDs round-robin (XCD =
// block id mod 8), each XCD with its own L2.  So block x goes to XCD
// x mod 8, which walks a contiguous stretch of a sequence in which every
// chunk of W tiles is followed by the chunk D*W ~ delta tiles later: the
// second read of an input region then comes from the same XCD's L2
// while the first is still resident.  tile_of() is a bijection between
// the blocks [0, 8P) that are not idle and the tiles [0, T).
struct TileOrder {
    u64 T, W, C, D, P;
};
GF2T_HD TileOrder tile_order(u64 T, u64 delta_pairs, u64 pairs_per_tile) {
    TileOrder o;
    o.T = T;
    o.W = 16;
    o.C = (T + o.W - 1) / o.W;
    u64 dt = (delta_pairs + pairs_per_tile / 2) / pairs_per_tile;   // tiles
    u64 D = (dt + o.W / 2) / o.W;                                     // chunks
    o.D = D ? D : 1;
    o.P = (o.C * o.W + 7) / 8;
    return o;
}
GF2T_HD int64_t tile_of(const TileOrder &o, u64 x) {
    if (x >= 8 * o.P) return -1;               // grid wider than this residue needs
    u64 pos = (x % 8) * o.P + x / 8;
    if (pos >= o.C * o.W) return -1;
    u64 cpos = pos / o.W, j = pos % o.W;
    u64 lo = o.C / o.D, big = o.C % o.D;   // big classes hold lo + 1 chunks
    u64 rho, q;
    if (cpos < big * (lo + 1)) {
        rho = cpos / (lo + 1);
        q = cpos % (lo + 1);
    } else {
        u64 c2 = cpos - big * (lo + 1);
        rho = big + c2 / lo;
        q = c2 % lo;
    }
    u64 tile = (rho + q * o.D) * o.W + j;
    return tile < o.T ? (int64_t)tile : -1;
}
GF2T_HD u64 sq_delta_pairs(const SqConst &c) {
    return c.w2.off() >= c.w3.off() ? c.w2.off() - c.w3.off() : 0;
}

// 64-bit window of p (pw words) at bit pos, restricted to bits [lo, hi)
GF2T_HD u64 pwin64(const u64 *p, size_t pw, u64 pos, u64 lo, u64 hi) {
    if (pos >= hi || pos + 64 <= lo) return 0;
    u64 v = win64(p, pw, pos);
    int lc = lo > pos ? (int)(lo - pos) : 0;
    int hc = (hi - pos) >= 64 ? 64 : (int)(hi - pos);
    return v & range_mask64(lc, hc);
}

// output word w of P mod x^r + x^s + 1 for a product P of L <= 2r-1 bits
GF2T_HD u64 fold_word(const u64 *p, size_t pw, u64 L, u64 r, u64 s, size_t w) {
    u64 b = (u64)w << 6;
    u64 lim1 = L < r ? L : r;
    u64 v = pwin64(p, pw, b, 0, lim1);
    if (L > r) {
        u64 t3 = L < 2 * r - s ? L : 2 * r - s;
        v ^= pwin64(p, pw, b + r, r, L);
        v ^= pwin64(p, pw, b + r - s, r, t3);
        v ^= pwin64(p, pw, b + 2 * r - s, 2 * r - s, L);
        v ^= pwin64(p, pw, b + 2 * r - 2 * s, 2 * r - s, L);
    }
    return v;
}

// ------------------------------------------------ sieve number theory
// x mod M for M = 2^d - 1 (Mersenne folding)
GF2T_HD u64 mod_mersenne(u64 x, int d, u64 M) {
    while (x > M) x = (x & M) + (x >> d);
    return x == M ? 0 : x;
}
// j is strictly below all its other rotations within d bits (d <= 32,
// M = 2^d - 1): a binary Lyndon word read most significant bit first,
// i.e. the representative of its class {j 2^i mod M} that lies in no
// proper subfield.  For d >= 2 such j is odd and below 2^(d-1).
GF2T_HD bool is_lyndon32(u32 j, int d, u32 M) {
    u32 jj = j;
    for (int i = 1; i < d; i++) {
        jj = ((jj << 1) | (jj >> (d - 1))) & M;
        if (jj <= j) return false;
    }
    return true;
}

GF2T_HD u32 gcd32(u32 a, u32 b) {
    while (b) { u32 t = a % b; a = b; b = t; }
    return a;
}
// inverse of a modulo m (gcd(a, m) = 1, m >= 2), extended Euclid in u32
GF2T_HD u32 inv32(u32 a, u32 m) {
    u32 r0 = m, r1 = a % m;
    long long t0 = 0, t1 = 1;
    while (r1) {
        u32 q = r0 / r1;
        u32 r2 = r0 - q * r1;
        long long t2 = t0 - (long long)q * t1;
        r0 = r1; r1 = r2; t0 = t1; t1 = t2;
    }
    if (t0 < 0) t0 += m;
    return (u32)t0;
}

// Minimal polynomial over GF(2) of beta in GF(2^d) (beta in no proper
// subfield) from the GF(2)-valued sequence a_k = Tr(beta^k): it obeys the
// recurrence whose characteristic polynomial is minpoly(beta) and has
// linear complexity d, so Berlekamp-Massey on a_0..a_{2d-1} returns the
// reversed minimal polynomial.  Tr(y) = parity(y & tmask) with tmask bit
// i = Tr(x^i).  2d field multiplies instead of the d^2 of the product
// over the conjugates; the polynomial is the same (it is unique).
// Returns the (d+1)-bit mask, or 0 if the complexity is not d.
GF2T_HD u64 minpoly_bm(u32 beta, u64 q, int d, u32 tmask) {
    u64 seq = 0;
    u32 pw = 1;
    for (int k = 0; k < 2 * d; k++) {
        seq |= (u64)(__builtin_popcount(pw & tmask) & 1) << k;
        pw = fmul(pw, beta, q, d);
    }
    u64 C = 1, B = 1, W = 0;
    int L = 0, m = 1;
    for (int n = 0; n < 2 * d; n++) {
        W = (W << 1) | ((seq >> n) & 1);           // bit j = a_{n-j}
        if (!(__builtin_popcountll(C & W) & 1)) { m++; continue; }
        u64 T = C;
        C ^= B << m;
        if (2 * L <= n) { L = n + 1 - L; B = T; m = 1; }
        else m++;
    }
    if (L != d) return 0;
    u64 mask = 0;
    for (int k = 0; k <= d; k++) mask |= ((C >> (d - k)) & 1) << k;
    return mask;
}
