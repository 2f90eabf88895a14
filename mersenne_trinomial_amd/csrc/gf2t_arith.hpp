// gf2t_arith.hpp -- the carry-less arithmetic the gfx950 kernels run,
// written once for device and host: the kernels include it through
// gf2t_internal.hpp, and a host build of the same functions is checked
// against the oracle on the CPU (tests/test_kernel_logic.py).
#pragma once
#include <stdint.h>
#include <stddef.h>

#if defined(__HIP__) || defined(__HIPCC__)
#define GF2T_HD __host__ __device__ __forceinline__
#else
#define GF2T_HD static inline
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


// ------------------------------------------------ GF(2^d), 32 < d <= 40
// 64 x 64 -> 128-bit carry-less product (Karatsuba over clmul32)
GF2T_HD void clmul64(u64 a, u64 b, u64 &lo, u64 &hi) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 l = clmul32(a0, b0), h = clmul32(a1, b1);
    u64 mid = clmul32(a0 ^ a1, b0 ^ b1) ^ l ^ h;
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
// 128-bit window of the virtual spread SP of `in` (SP[2k] = bit k of in,
// odd positions 0) at SP position P, keeping only SP indices in [lo, hi).
GF2T_HD void vspread128(const u64 *in, size_t nw, u64 P, u64 lo, u64 hi, u64 &w0,
                        u64 &w1) {
    if (P >= hi || P + 128 <= lo) { w0 = 0; w1 = 0; return; }
    u64 v = win64(in, nw, (P + 1) >> 1);
    u64 a = spread32((u32)v), b = spread32((u32)(v >> 32));
    if (P & 1) { a <<= 1; b <<= 1; }
    int lc = lo > P ? (int)(lo - P) : 0;
    int hc = (hi - P) >= 128 ? 128 : (int)(hi - P);
    w0 = a & range_mask64(lc, hc);
    w1 = b & range_mask64(lc - 64, hc - 64);
}

// output words 2i, 2i+1 of in^2 mod x^r + x^s + 1:
//   SP[0,r) + SP[r,2r-1)>>r + SP[r,2r-s)>>(r-s)
//   + SP[2r-s,2r-1)>>(2r-s) + SP[2r-s,2r-1)>>(2r-2s)
GF2T_HD void square_fold_pair(const u64 *in, size_t nw, u64 r, u64 s, size_t i, u64 &o0,
                              u64 &o1) {
    const u64 b = (u64)i << 7;
    u64 t0, t1;
    vspread128(in, nw, b, 0, r, o0, o1);
    vspread128(in, nw, b + r, r, 2 * r - 1, t0, t1);
    o0 ^= t0; o1 ^= t1;
    vspread128(in, nw, b + r - s, r, 2 * r - s, t0, t1);
    o0 ^= t0; o1 ^= t1;
    vspread128(in, nw, b + 2 * r - s, 2 * r - s, 2 * r - 1, t0, t1);
    o0 ^= t0; o1 ^= t1;
    vspread128(in, nw, b + 2 * r - 2 * s, 2 * r - s, 2 * r - 1, t0, t1);
    o0 ^= t0; o1 ^= t1;
}

// ---- fast path: unmasked windows, 32-bit spreads
// spread of the low 16 bits of x into 32 bits (bit k -> bit 2k)
GF2T_HD u32 spread16(u32 x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
GF2T_HD void spread64(u64 v, u64 &lo, u64 &hi) {
    u32 a = (u32)v, b = (u32)(v >> 32);
    lo = (u64)spread16(a) | ((u64)spread16(a >> 16) << 32);
    hi = (u64)spread16(b) | ((u64)spread16(b >> 16) << 32);
}

// Window k of the squaring reads the 64 input bits starting at input bit
// 64i + Q_k (i = output pair index) and, when the spread position is odd,
// shifts the spread left by one.  Q_k, and so the word offset, the bit
// shift and the parity, are the same for every pair of one residue.
struct SqWin {
    u64 off;
    u32 sh, odd;
};
struct SqConst {
    u64 r, s;
    size_t nw;
    SqWin w2, w3, w4, w5;
};
GF2T_HD SqWin sq_win(u64 P0) {   // P0 = window offset in spread positions (b = 0)
    u64 Q = (P0 + 1) >> 1;
    SqWin w;
    w.off = Q >> 6;
    w.sh = (u32)(Q & 63);
    w.odd = (u32)(P0 & 1);
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
GF2T_HD void win_spread(const u64 *in, size_t nw, size_t i, const SqWin &w, u64 &lo, u64 &hi) {
    size_t k = i + (size_t)w.off;
    u64 v = k < nw ? in[k] : 0;
    if (w.sh) {
        u64 h = k + 1 < nw ? in[k + 1] : 0;
        v = (v >> w.sh) | (h << (64 - w.sh));
    }
    spread64(v, lo, hi);
    if (w.odd) { lo <<= 1; hi <<= 1; }   // spread bits are even: no carry between words
}

// Output pair i of in^2 mod T.  Away from the points s, 2s and r every
// window is either fully inside or fully outside its clip range:
//   [0, s)  : SP[0,r) + SP[r,..)>>r + SP[2r-s,..)>>(2r-s)          (W1 W2 W4)
//   [s, 2s) : W1 + W2 + SP[r,2r-s)>>(r-s) + SP[2r-s,..)>>(2r-2s)   (W1 W2 W3 W5)
//   [2s, r) : W1 + W2 + W3
// so those pairs skip all masking; pairs straddling s, 2s or r take the
// clipped general path.
GF2T_HD void square_fold_pair_fast(const u64 *in, const SqConst &C, size_t i, u64 &o0, u64 &o1) {
    const u64 b = (u64)i << 7;
    const u64 s = C.s, r = C.r;
    int region;
    if (b + 128 <= s) region = 0;
    else if (b >= s && b + 128 <= 2 * s) region = 1;
    else if (b >= 2 * s && b + 128 <= r) region = 2;
    else region = 3;
    if (region == 3) {
        square_fold_pair(in, C.nw, r, s, i, o0, o1);
        return;
    }
    u64 t0, t1;
    u64 x = i < C.nw ? in[i] : 0;
    spread64(x, o0, o1);
    win_spread(in, C.nw, i, C.w2, t0, t1);
    o0 ^= t0; o1 ^= t1;
    if (region == 0) {
        win_spread(in, C.nw, i, C.w4, t0, t1);
        o0 ^= t0; o1 ^= t1;
    } else {
        win_spread(in, C.nw, i, C.w3, t0, t1);
        o0 ^= t0; o1 ^= t1;
        if (region == 1) {
            win_spread(in, C.nw, i, C.w5, t0, t1);
            o0 ^= t0; o1 ^= t1;
        }
    }
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
    return c.w2.off >= c.w3.off ? c.w2.off - c.w3.off : 0;
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
