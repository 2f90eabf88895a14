// Host build of the kernels' arithmetic (mersenne_trinomial_amd/csrc/
// gf2t_arith.hpp) for CPU checks against the oracle.  Test code only.
#include "gf2t_arith.hpp"

extern "C" {
void kl_square_fold(const u64 *in, u64 *out, u64 r, u64 s) {
    size_t nw = (size_t)((r + 63) >> 6);
    for (size_t i = 0; i < (nw + 1) / 2; i++) {
        u64 o0, o1;
        square_fold_pair(in, nw, r, s, i, o0, o1);
        out[2 * i] = o0;
        if (2 * i + 1 < nw) out[2 * i + 1] = o1;
    }
}
void kl_square_fold_fast(const u64 *in, u64 *out, u64 r, u64 s) {
    SqConst c = sq_const(r, s);
    for (size_t i = 0; i < (c.nw + 1) / 2; i++) {
        u64 o0, o1;
        square_fold_pair_fast(in, c, i, o0, o1);
        out[2 * i] = o0;
        if (2 * i + 1 < c.nw) out[2 * i + 1] = o1;
    }
}
void kl_fold(const u64 *p, u64 L, u64 *out, u64 r, u64 s) {
    size_t nw = (size_t)((r + 63) >> 6), pw = (size_t)((L + 63) >> 6);
    for (size_t w = 0; w < nw; w++) out[w] = fold_word(p, pw, L, r, s, w);
}
u64 kl_gf64_mul(u64 a, u64 b) { return gf64_mul(a, b); }
u64 kl_clmul32(u32 a, u32 b) { return clmul32(a, b); }
u32 kl_fmul(u32 a, u32 b, u64 q, int d) { return fmul(a, b, q, d); }
u64 kl_range_mask64(int lc, int hc) { return range_mask64(lc, hc); }
int64_t kl_tile_of(u64 T, u64 delta_pairs, u64 ppt, u64 x, u64 *nblocks) {
    TileOrder o = tile_order(T, delta_pairs, ppt);
    *nblocks = 8 * o.P;
    return tile_of(o, x);
}
u32 kl_gcd32(u32 a, u32 b) { return gcd32(a, b); }
u32 kl_inv32(u32 a, u32 m) { return inv32(a, m); }
u64 kl_mod_mersenne(u64 x, int d) { return mod_mersenne(x, d, ((u64)1 << d) - 1); }
u64 kl_minpoly_bm(u32 beta, u64 q, int d, u32 tmask) { return minpoly_bm(beta, q, d, tmask); }
u64 kl_fmul64(u64 a, u64 b, u64 q, int d) { return fmul64(a, b, q, d); }
u64 kl_minpoly_bm64(u64 beta, u64 q, int d, u64 tmask) { return minpoly_bm64(beta, q, d, tmask); }
u64 kl_mulmod64(u64 a, u64 b, u64 m) { return mulmod64(a, b, m); }
u64 kl_inv64(u64 a, u64 m) { return inv64(a, m); }
u64 kl_gcd64(u64 a, u64 b) { return gcd64(a, b); }
u64 kl_count_classes64(int d) {
    u64 M = ((u64)1 << d) - 1, n = 0;
    for (u64 c = 0; c < ((u64)1 << (d - 2)); c++) n += is_lyndon64(2 * c + 1, d, M);
    return n;
}
// the sieve's class enumeration: candidates j = 2c + 1, c < 2^(d-2)
u64 kl_count_classes(int d) {
    u32 M = (u32)(((u64)1 << d) - 1);
    u64 n = 0;
    for (u64 c = 0; c < ((u64)1 << (d - 2)); c++) n += is_lyndon32((u32)(2 * c + 1), d, M);
    return n;
}
// the same over every j in [1, M) (no candidate restriction)
u64 kl_count_classes_all(int d) {
    u32 M = (u32)(((u64)1 << d) - 1);
    u64 n = 0;
    for (u64 j = 1; j < M; j++) n += is_lyndon32((u32)j, d, M);
    return n;
}
}
