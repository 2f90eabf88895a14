// Host build of the kernels' arithmetic (mersenne_trinomial_amd/csrc/
// gf2t_arith.hpp) for CPU checks against the oracle.  Test code only.
#include "gf2t_arith.hpp"
#include <stdlib.h>
#include <vector>

// the two-squaring tile's window copies, with every read checked
// against the words the G phase wrote: holds[x] is the G word that LDS
// word x was last given (~0: none)
struct CheckedTile {
    Sq2Tile t;
    const std::vector<u64> *holds;
    u64 operator()(int k, size_t w) const {
        const u64 at = w - t.base[k];
        if (w < t.base[k] || at >= SQ2_LEN || (*holds)[t.at[k] + at] != w) abort();
        return t(k, w);
    }
    u32 ld32(int k, u32 q) const {
        const u64 w = (*this)(k, q >> 1);
        return (q & 1) ? (u32)(w >> 32) : (u32)w;
    }
    void raw3(int k, u32 q, u32 sh, u32 &a, u32 &b, u32 &c) const {
        if (sh > 31) abort();
        a = ld32(k, q); b = ld32(k, q + 1); c = ld32(k, q + 2);
    }
};

// the kernels' global source with every 32-bit read checked against the
// 2 nw words of the residue; w12 as in SqArrayT<true> (word q + 2 read
// for every window), used where sq_w12_ok(r) lets the kernels use it
struct CheckedArray {
    const u64 *p;
    size_t nw;
    bool w12;
    u64 operator()(int, size_t k) const { return k < nw ? p[k] : 0; }
    u32 w32(u32 q) const {
        if ((size_t)q >= 2 * nw) abort();
        return (q & 1) ? (u32)(p[q >> 1] >> 32) : (u32)p[q >> 1];
    }
    void raw3(int, u32 q, u32 sh, u32 &a, u32 &b, u32 &c) const {
        if (sh > 31) abort();
        a = w32(q);
        b = w32(q + 1);
        c = w32(w12 ? q + 2 : q + 1 + (sh != 0));
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
        square_fold_pair_fast_src(CheckedArray{in, c.nw, sq_w12_ok(r)}, c, i, o0, o1);
        out[2 * i] = o0;
        if (2 * i + 1 < c.nw) out[2 * i + 1] = o1;
    }
}
// the G phase of one tile: the windows of pl.compute, one G pair per
// thread, into the plan's slots
static void g_phase(const CheckedArray &in, const SqConst &c, const Sq2Plan &pl,
                    std::vector<u64> &gw,
                    std::vector<u64> &holds, u64 &formed) {
    for (int k = 0; k < SQ2_WINS; k++) {
        if (!((pl.compute >> k) & 1)) continue;
        formed++;
        const u64 p0 = pl.base[k] >> 1;
        const int greg = sq_tile_region(c, p0, p0 + SQ2_GP);
        for (u32 t = 0; t < SQ2_GP; t++) {
            u64 o0, o1;
            if (greg < 3) {
                SqRawPair w = sq_raw_pair(in, c, (u32)(p0 + t), greg);
                sq_combine_pair(w, c, greg, o0, o1);
            } else {
                sq2_g_pair_clipped(in, c, p0 + t, o0, o1);
            }
            const size_t at = pl.at[k] + 2 * t;
            gw[at] = o0;
            gw[at + 1] = o1;
            holds[at] = 2 * (p0 + t);
            holds[at + 1] = 2 * (p0 + t) + 1;
        }
    }
}
static void h_phase(const SqConst &c, size_t i0, size_t i1, int treg, const Sq2Plan &pl,
                    const std::vector<u64> &gw, const std::vector<u64> &holds, u64 *out,
                    std::vector<unsigned char> &hits) {
    CheckedTile src{sq2_tile(gw.data(), pl), &holds};
    for (size_t i = i0; i < i1; i++) {
        u64 o0, o1;
        if (treg < 3) sq_pair_region(src, c, (u32)i, treg, o0, o1);
        else square_fold_pair_fast_src(src, c, i, o0, o1);
        out[2 * i] = o0;
        if (2 * i + 1 < c.nw) out[2 * i + 1] = o1;
        hits[i]++;
    }
}
// the two-squaring kernel's plain tiles, run one after another
void kl_square2_fold(const u64 *in, u64 *out, u64 r, u64 s) {
    SqConst c = sq_const(r, s);
    const size_t np = (c.nw + 1) / 2;
    std::vector<u64> gw((size_t)SQ2_WINS * SQ2_LEN), holds(gw.size());
    std::vector<unsigned char> hits(np);
    u64 formed = 0;
    for (size_t i0 = 0; i0 < np; i0 += SQ2_TP) {
        const size_t i1 = i0 + SQ2_TP < np ? i0 + SQ2_TP : np;
        const int treg = sq_tile_region(c, i0, i1);
        const Sq2Plan pl = sq2_plan(c, i0, sq2_need(treg));
        std::fill(holds.begin(), holds.end(), ~0ULL);
        g_phase(CheckedArray{in, c.nw, sq_w12_ok(r)}, c, pl, gw, holds, formed);
        h_phase(c, i0, i1, treg, pl, gw, holds, out, hits);
    }
}
// the chained work units of k_square2_fold_batch, each a fresh block
// (LDS of unknown content) walking its tiles; every output pair must be
// formed exactly once.  stats: units with tiles, tiles, windows formed.
void kl_square2_fold_units(const u64 *in, u64 *out, u64 r, u64 s, u64 *stats) {
    SqConst c = sq_const(r, s);
    const Sq2Chain h = sq2_chain(c);
    std::vector<u64> gw((size_t)SQ2_WINS * SQ2_LEN), holds(gw.size());
    std::vector<unsigned char> hits(h.np);
    u64 units = 0, tiles = 0, formed = 0;
    const u64 nu = sq2_units(h);
    for (u64 u = 0; u < nu; u++) {
        Sq2Unit w;
        if (!sq2_unit(h, (u32)u, w)) continue;
        if (w.width == 0 || w.width > (h.chained ? SQ2_CW : SQ2_TP)) abort();
        units++;
        std::fill(gw.begin(), gw.end(), 0x5A5A5A5A5A5A5A5AULL);
        std::fill(holds.begin(), holds.end(), ~0ULL);
        u32 prev_need = 0, prev_base4 = 0;
        for (u32 j = w.j0; j < w.j1; j++) {
            const u32 i0 = w.start + j * h.period;
            if (i0 >= h.np) abort();
            const u32 i1 = i0 + w.width < h.np ? i0 + w.width : h.np;
            // the H phase's 512 pair slots start at i0 rounded down to 8
            if ((i0 & 7) + (i1 - i0) > 512) abort();
            const int treg = sq_tile_region(c, i0, i1);
            const u32 need = sq2_need(treg);
            const Sq2Plan pl = sq2_chain_plan(c, h, i0, w.width, need, j - w.j0, prev_need, prev_base4);
            if ((pl.compute & ~need) || !(pl.compute & 2u)) abort();
            g_phase(CheckedArray{in, c.nw, sq_w12_ok(r)}, c, pl, gw, holds, formed);
            h_phase(c, i0, i1, treg, pl, gw, holds, out, hits);
            tiles++;
            prev_need = need;
            prev_base4 = pl.base[3];
        }
    }
    for (size_t i = 0; i < h.np; i++)
        if (hits[i] != 1) abort();
    stats[0] = units;
    stats[1] = tiles;
    stats[2] = formed;
}
// the single-squaring batch kernel's per-tile region path
void kl_square_fold_tiles(const u64 *in, u64 *out, u64 r, u64 s, u32 tile_pairs) {
    SqConst c = sq_const(r, s);
    const size_t np = (c.nw + 1) / 2;
    const CheckedArray src{in, c.nw, sq_w12_ok(r)};
    for (size_t t0 = 0; t0 < np; t0 += tile_pairs) {
        const size_t t1 = t0 + tile_pairs < np ? t0 + tile_pairs : np;
        const int treg = sq_tile_region(c, t0, t1);
        for (size_t i = t0; i < t1; i++) {
            u64 o0, o1;
            if (treg < 3) sq_pair_region(src, c, (u32)i, treg, o0, o1);
            else square_fold_pair_fast_src(src, c, i, o0, o1);
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
