// Host build of the kernels' arithmetic (mersenne_trinomial_amd/csrc/
// gf2t_arith.hpp) for CPU checks against the oracle.  Test code only.
#include "gf2t_arith.hpp"
#include <stdlib.h>
#include <vector>

// the two-squaring tile's window copies, with every read checked
// against the words the G phase wrote
struct CheckedTile {
    Sq2Tile t;
    const std::vector<char> *written;
    u64 operator()(int k, size_t w) const {
        const u64 at = w - t.base[k];
        if (w < t.base[k] || at >= SQ2_LEN || !(*written)[(size_t)k * SQ2_LEN + at]) abort();
        return t(k, w);
    }
    u32 ld32(int k, u32 q) const {
        const u64 w = (*this)(k, q >> 1);
        return (q & 1) ? (u32)(w >> 32) : (u32)w;
    }
    void raw3(int k, u32 q, u32 sh, u32 &a, u32 &b, u32 &c) const {
        if (sh > 31) abort();
        a = ld32(k, q); b = ld32(k, q + 1); c = ld32(k, q + 1 + (sh != 0));
    }
};

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
// the two-squaring kernel's tiles, run one after another: the G phase
// (one G pair per thread and needed window), then the tile's H pairs
// through the same calls the kernel makes
void kl_square2_fold(const u64 *in, u64 *out, u64 r, u64 s) {
    SqConst c = sq_const(r, s);
    const size_t np = (c.nw + 1) / 2;
    std::vector<u64> gw((size_t)SQ2_WINS * SQ2_LEN);
    std::vector<char> written(gw.size());
    for (size_t i0 = 0; i0 < np; i0 += SQ2_TP) {
        const size_t i1 = i0 + SQ2_TP < np ? i0 + SQ2_TP : np;
        const int treg = sq_tile_region(c, i0, i1);
        const u32 need = sq2_need(treg);
        std::fill(gw.begin(), gw.end(), 0x5A5A5A5A5A5A5A5AULL);
        std::fill(written.begin(), written.end(), 0);
        for (int k = 0; k < SQ2_WINS; k++) {
            if (!((need >> k) & 1)) continue;
            const u64 p0 = sq2_p0(c, i0, k);
            const int greg = sq_tile_region(c, p0, p0 + SQ2_GP);
            for (u32 t = 0; t < SQ2_GP; t++) {
                u64 o0, o1;
                if (greg < 3) {
                    SqRawPair w = sq_raw_pair(SqArray{in, c.nw}, c, (u32)(p0 + t), greg);
                    sq_combine_pair(w, c, greg, o0, o1);
                } else {
                    sq2_g_pair_clipped(in, c, p0 + t, o0, o1);
                }
                gw[(size_t)k * SQ2_LEN + 2 * t] = o0;
                gw[(size_t)k * SQ2_LEN + 2 * t + 1] = o1;
                written[(size_t)k * SQ2_LEN + 2 * t] = written[(size_t)k * SQ2_LEN + 2 * t + 1] = 1;
            }
        }
        CheckedTile src{sq2_tile(gw.data(), c, i0), &written};
        for (size_t i = i0; i < i1; i++) {
            u64 o0, o1;
            if (treg < 3) sq_pair_region(src, c, (u32)i, treg, o0, o1);
            else square_fold_pair_fast_src(src, c, i, o0, o1);
            out[2 * i] = o0;
            if (2 * i + 1 < c.nw) out[2 * i + 1] = o1;
        }
    }
}
// the single-squaring batch kernel's per-tile region path
void kl_square_fold_tiles(const u64 *in, u64 *out, u64 r, u64 s, u32 tile_pairs) {
    SqConst c = sq_const(r, s);
    const size_t np = (c.nw + 1) / 2;
    const SqArray src{in, c.nw};
    for (size_t t0 = 0; t0 < np; t0 += tile_pairs) {
        const size_t t1 = t0 + tile_pairs < np ? t0 + tile_pairs : np;
        const int treg = sq_tile_region(c, t0, t1);
        for (size_t i = t0; i < t1; i++) {
            u64 o0, o1;
            if (treg < 3) sq_pair_region(src, c, (u32)i, treg, o0, o1);
            else square_fold_pair_fast(in, c, i, o0, o1);
            out[2 * i] = o0;
            if (2 * i + 1 < c.nw) out[2 * i + 1] = o1;
        }
    }
}
u64 kl_spread32p(u32 x) { return spread32p(x); }
u64 kl_interleave32(u32 e, u32 o) { return interleave32(e, o); }
u64 kl_spread32(u32 x) { return spread32(x); }
u64 kl_funnel64(u64 a, u64 b, u32 sh) { return funnel64(a, b, sh); }
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
