// square.hip -- (b) squaring + trinomial fold on gfx950.
//
// F <- F^2 mod T, T = x^r + x^s + 1, 1 <= s, 2s < r.  Squaring over GF(2)
// is the bit spread SP (SP[2k] = F bit k, SP[odd] = 0); the fold
// x^e -> x^(e-r) + x^(e-r+s) applied twice reduces SP (degree <= 2r-2)
// to r bits.  Writing SP = L + x^r H, H = H_lo + x^(r-s) H_hi:
//   F^2 mod T = SP[0,r) + SP[r,2r-1)>>r + SP[r,2r-s)>>(r-s)
//             + SP[2r-s,2r-1)>>(2r-s) + SP[2r-s,2r-1)>>(2r-2s)
// (the same five windows as the reference's fused kernel,
// src/tsfactor.cu:128-186).  Each output word is written by exactly one
// lane (gather form), so there are no write races.
//
// Mapping: one lane produces one 128-bit output pair (two words) from
// 64-bit input windows; waves cover 64 consecutive pairs = 1 KiB of
// output, so stores and the dominant loads are fully coalesced.  Per
// residue-step the kernel moves 16*nw algorithmic bytes (read F, write
// F^2); the second read of the upper input half by window 3 hits in
// L2 / Infinity Cache.
#include "gf2t_internal.hpp"
#include "host_poly.hpp"
#include "../../include/gf2t.h"
#include <vector>
#include <algorithm>
#include <numeric>
#include <cstring>
#include <cstdlib>

namespace {

constexpr int SQ_THREADS = 256;

// W12 (template argument of the squaring kernels): the window words are
// read as one 12-byte load; the host sets it when no residue of the launch
// has r a multiple of 64 (sq_w12_ok)
template <bool W12>
__global__ __launch_bounds__(SQ_THREADS) void k_square_fold(
        const u64 *__restrict__ in, u64 *__restrict__ out, const SqConst c) {
    size_t i = (size_t)blockIdx.x * SQ_THREADS + threadIdx.x;
    if (i >= (c.nw + 1) / 2) return;
    u64 o0, o1;
    square_fold_pair_fast<W12>(in, c, i, o0, o1);
    out[2 * i] = o0;
    if (2 * i + 1 < c.nw) out[2 * i + 1] = o1;
}

struct SqDesc {
    const u64 *in;
    u64 *out;
    SqConst c;
    const int *order;   // block -> tile (-1 idle), or null for identity
    Sq2Chain h;         // two-squaring tile layout
};

// output pair i (one 16-byte store when both words exist)
template <bool NT>
__device__ __forceinline__ void store_pair(u64 *out_, size_t nw, size_t i, u64 o0, u64 o1) {
    typedef u64 v2u __attribute__((ext_vector_type(2)));
    GF2T_GLOBAL u64 *out = (GF2T_GLOBAL u64 *)out_;
    if (2 * i + 1 < nw) {
        GF2T_GLOBAL v2u *p = (GF2T_GLOBAL v2u *)(out + 2 * i);
        if (NT) __builtin_nontemporal_store(v2u{o0, o1}, p);
        else *p = v2u{o0, o1};
    } else {
        if (NT) __builtin_nontemporal_store(o0, out + 2 * i);
        else out[2 * i] = o0;
    }
}

// batched: blockIdx.y selects the residue, blockIdx.x a tile of
// SQ_THREADS*PPT output pairs (pairs j*SQ_THREADS + lane for j < PPT)
template <int PPT, bool NT, bool W12>
__global__ __launch_bounds__(SQ_THREADS) void k_square_fold_batch(
        const SqDesc *__restrict__ desc) {
    const SqDesc d = desc[blockIdx.y];
    int t = (int)blockIdx.x;
    if (d.order) {
        t = __builtin_amdgcn_readfirstlane(d.order[blockIdx.x]);   // one value per block
        if (t < 0) return;
    }
    const size_t np = (d.c.nw + 1) / 2;
    const size_t t0 = (size_t)t * SQ_THREADS * PPT;
    if (t0 >= np) return;
    // one region for the whole tile (all but a few tiles per residue)
    const int treg = sq_tile_region(d.c, t0, t0 + SQ_THREADS * PPT < np ? t0 + SQ_THREADS * PPT : np);
    const SqArrayT<W12> src{d.in, d.c.nw};
    // streaming stores: F^2 is not re-read before the next launch (one
    // 16-byte store per pair; residue buffers are 16-byte aligned)
    if (treg < 3) {
        sq_variant(treg, d.c, [&](auto v) {
#pragma unroll
            for (int j = 0; j < PPT; j++) {
                const size_t i = t0 + (size_t)j * SQ_THREADS + threadIdx.x;
                if (i < np) {
                    u64 o0, o1;
                    v.pair(src, d.c, (u32)i, o0, o1);
                    store_pair<NT>(d.out, d.c.nw, i, o0, o1);
                }
            }
        });
    } else {
        for (int j = 0; j < PPT; j++) {
            const size_t i = t0 + (size_t)j * SQ_THREADS + threadIdx.x;
            if (i < np) {
                u64 o0, o1;
                square_fold_pair_fast<W12>(d.in, d.c, i, o0, o1);
                store_pair<NT>(d.out, d.c.nw, i, o0, o1);
            }
        }
    }
}

// ---- two squarings per launch: H = (F^2)^2 mod T
// One tile, the output pairs [i0, i1) (at most SQ2_TP): the G phase forms
// the windows of pl.compute, one G pair per thread and window, into the
// plan's LDS slots; the H phase then reads G through the window copies of
// the plan.  LOADFIRST issues the loads of every window before any is
// combined (a stand-alone tile); without it only windows 1 and 2, which
// every tile forms, go first and the others are loaded as they are
// formed, which keeps the work-unit loop at 56 VGPRs (7 waves per SIMD)
// where the full load-first form needed 84.
template <bool NT, bool W12, bool LOADFIRST = true>
__device__ __forceinline__ void square2_run(const u64 *__restrict__ in, u64 *__restrict__ out,
                                            const SqConst &c, size_t i0, size_t i1, int treg,
                                            const Sq2Plan &pl, u64 *gw) {
    static_assert(SQ_THREADS == SQ2_GP && 2 * SQ_THREADS >= SQ2_TP + 7, "tile shape");
    typedef u64 v2u __attribute__((ext_vector_type(2)));
    const SqArrayT<W12> fsrc{in, c.nw};
    int greg[SQ2_WINS];
    SqRawPair raw[SQ2_WINS];
#pragma unroll
    for (int k = 0; k < SQ2_WINS; k++) {
        const u64 p0 = pl.base[k] >> 1;
        greg[k] = ((pl.compute >> k) & 1) ? sq_tile_region(c, p0, p0 + SQ2_GP) : -1;
    }
    // loads issued ahead of any combining: every window's (LOADFIRST), or
    // those of windows 1 and 2, which every tile forms
#pragma unroll
    for (int k = 0; k < (LOADFIRST ? SQ2_WINS : 2); k++)
        if (greg[k] >= 0 && greg[k] < 3)
            raw[k] = sq_raw_pair(fsrc, c, (u32)((pl.base[k] >> 1) + threadIdx.x), greg[k]);
#pragma unroll
    for (int k = 0; k < SQ2_WINS; k++) {
        if (greg[k] < 0) continue;
        u64 o0, o1;
        if (greg[k] < 3) {
            if (!LOADFIRST && k >= 2)
                raw[k] = sq_raw_pair(fsrc, c, (u32)((pl.base[k] >> 1) + threadIdx.x), greg[k]);
            sq_combine_pair(raw[k], c, greg[k], o0, o1);
        } else {
            sq2_g_pair_clipped(fsrc, c, (pl.base[k] >> 1) + threadIdx.x, o0, o1);
        }
        *(v2u *)(gw + pl.at[k] + 2 * threadIdx.x) = v2u{o0, o1};
    }
    __syncthreads();
    const Sq2Tile src = sq2_tile(gw, pl);
    // pair slots from i0 rounded down to a 128-byte line: each wave stores
    // whole lines (non-temporal stores of part of a line go to memory
    // as partial writes); the slots below i0 idle
    const size_t ib = i0 & ~(size_t)7;
    if (treg < 3) {
        sq_variant(treg, c, [&](auto v) {
            for (int j = 0; j < 2; j++) {
                const size_t i = ib + (size_t)j * SQ_THREADS + threadIdx.x;
                if (i >= i0 && i < i1) {
                    u64 o0, o1;
                    v.pair(src, c, (u32)i, o0, o1);
                    store_pair<NT>(out, c.nw, i, o0, o1);
                }
            }
        });
    } else {
        for (int j = 0; j < 2; j++) {
            const size_t i = ib + (size_t)j * SQ_THREADS + threadIdx.x;
            if (i >= i0 && i < i1) {
                u64 o0, o1;
                square_fold_pair_fast_src(src, c, i, o0, o1);
                store_pair<NT>(out, c.nw, i, o0, o1);
            }
        }
    }
}
// the plain tile number `tile` (pairs from tile * SQ2_TP on)
template <bool NT, bool W12>
__device__ __forceinline__ void square2_tile(const u64 *__restrict__ in, u64 *__restrict__ out,
                                             const SqConst &c, size_t tile, u64 *gw) {
    const size_t np = (c.nw + 1) / 2;
    const size_t i0 = tile * SQ2_TP;
    if (i0 >= np) return;   // whole block
    const size_t i1 = i0 + SQ2_TP < np ? i0 + SQ2_TP : np;
    const int treg = sq_tile_region(c, i0, i1);
    square2_run<NT, W12, false>(in, out, c, i0, i1, treg, sq2_plan(c, i0, sq2_need(treg)), gw);
}
// work unit u: its tiles one after another, each taking windows 3 and 5
// from the copies of the tile before it where the plan says so
template <bool NT, bool W12>
__device__ __forceinline__ void square2_unit(const u64 *__restrict__ in, u64 *__restrict__ out,
                                             const SqConst &c, const Sq2Chain &h, u32 u,
                                             u64 *gw) {
    Sq2Unit w;
    if (!sq2_unit(h, u, w)) return;   // whole block
    u32 prev_need = 0, prev_base4 = 0;
    for (u32 j = w.j0; j < w.j1; j++) {
        if (j > w.j0) __syncthreads();   // the previous H phase is done with the slots
        const u32 i0 = w.start + j * h.period;
        const u32 i1 = i0 + w.width < h.np ? i0 + w.width : h.np;
        const int treg = sq_tile_region(c, i0, i1);
        const u32 need = sq2_need(treg);
        const Sq2Plan pl = sq2_chain_plan(c, h, i0, w.width, need, j - w.j0, prev_need, prev_base4);
        square2_run<NT, W12, false>(in, out, c, i0, i1, treg, pl, gw);
        prev_need = need;
        prev_base4 = pl.base[3];
    }
}

// batched: blockIdx.y selects the residue, blockIdx.x the work unit.
// Eight waves per SIMD (106 SGPRs would allow seven; 23 of them then live
// in VGPR lanes, no scratch): about 2% faster per launch.
template <bool NT, bool W12>
__global__ __launch_bounds__(SQ_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void
k_square2_fold_batch(
        const SqDesc *__restrict__ desc) {
    __shared__ __align__(16) u64 gw[SQ2_WINS * SQ2_LEN];
    const SqDesc d = desc[blockIdx.y];
    square2_unit<NT, W12>(d.in, d.out, d.c, d.h, blockIdx.x, gw);
}
template <bool W12>
__global__ __launch_bounds__(SQ_THREADS) void k_square2_fold(
        const u64 *__restrict__ in, u64 *__restrict__ out, const SqConst c, const Sq2Chain h) {
    __shared__ __align__(16) u64 gw[SQ2_WINS * SQ2_LEN];
    square2_unit<false, W12>(in, out, c, h, blockIdx.x, gw);
}

// slab of n residues (pitch words apart), residue y with s = svec[y / group],
// advanced by two squarings: the h-table step of the blocked DDF taken
// twice in one pass.  Tiles beyond the residue exit before any barrier.
__device__ __forceinline__ u64 uniform_u64(u64 v) {
    return ((u64)__builtin_amdgcn_readfirstlane((u32)(v >> 32)) << 32) |
           __builtin_amdgcn_readfirstlane((u32)v);
}
// One tile of one residue per block: y = blockIdx.z * gridDim.y +
// blockIdx.y (a loop over residues here held its whole state across
// iterations, at 160-210 VGPRs and 2-3 waves per SIMD, against 47 and 8
// without it).
template <bool W12>
__global__ __launch_bounds__(SQ_THREADS) void k_square2_fold_slab_g(
        const u64 *__restrict__ in, u64 *__restrict__ out, u64 pitch, u64 r,
        const u64 *__restrict__ svec, unsigned n, unsigned group) {
    __shared__ __align__(16) u64 gw[SQ2_WINS * SQ2_LEN];
    const unsigned y = blockIdx.z * gridDim.y + blockIdx.y;
    if (y >= n) return;   // whole block
    const SqConst c = sq_const(r, uniform_u64(svec[y / group]));
    square2_tile<false, W12>(in + (u64)y * pitch, out + (u64)y * pitch, c, blockIdx.x, gw);
}

// The same two squarings of a slab in work units (as the batched kernel):
// desc[g] holds group g's constants and tile layout, computed on the host
// once per cohort (gf2t_sq2_slab_plan); blockIdx.x is the work unit.
struct Sq2SlabDesc {
    SqConst c;
    Sq2Chain h;
};
template <bool NT, bool W12>
__global__ __launch_bounds__(SQ_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void
k_square2_fold_slab_u(const u64 *__restrict__ in, u64 *__restrict__ out, u64 pitch,
                      const Sq2SlabDesc *__restrict__ desc, unsigned n, unsigned group) {
    __shared__ __align__(16) u64 gw[SQ2_WINS * SQ2_LEN];
    const unsigned y = blockIdx.z * gridDim.y + blockIdx.y;
    if (y >= n) return;   // whole block
    const Sq2SlabDesc d = desc[y / group];
    square2_unit<NT, W12>(in + (u64)y * pitch, out + (u64)y * pitch, d.c, d.h, blockIdx.x, gw);
}

// slab: residue y at in + y*pitch, its s in svec[y]
__global__ __launch_bounds__(SQ_THREADS) void k_square_fold_slab(
        const u64 *__restrict__ in, u64 *__restrict__ out, u64 pitch, u64 r,
        const u64 *__restrict__ svec) {
    const SqConst c = sq_const(r, svec[blockIdx.y]);
    size_t i = (size_t)blockIdx.x * SQ_THREADS + threadIdx.x;
    if (i >= (c.nw + 1) / 2) return;
    in += (u64)blockIdx.y * pitch;
    out += (u64)blockIdx.y * pitch;
    u64 o0, o1;
    square_fold_pair_fast(in, c, i, o0, o1);
    out[2 * i] = o0;
    if (2 * i + 1 < c.nw) out[2 * i + 1] = o1;
}

__global__ __launch_bounds__(SQ_THREADS) void k_fold(
        const u64 *__restrict__ p, size_t pw, u64 L, u64 *__restrict__ out,
        size_t nw, u64 r, u64 s) {
    size_t w = (size_t)blockIdx.x * SQ_THREADS + threadIdx.x;
    if (w >= nw) return;
    out[w] = fold_word(p, pw, L, r, s, w);
}

// dst ^= x*src mod T
__global__ __launch_bounds__(SQ_THREADS) void k_mulx_xor(
        u64 *__restrict__ dst, const u64 *__restrict__ src, size_t nw, u64 r,
        u64 s) {
    size_t w = (size_t)blockIdx.x * SQ_THREADS + threadIdx.x;
    if (w >= nw) return;
    u64 a = src[w];
    u64 prev = w ? src[w - 1] : 0;
    u64 v = (a << 1) | (prev >> 63);
    if (w == (size_t)(r >> 6)) v &= ~(1ULL << (r & 63));
    u64 wrap = (src[(r - 1) >> 6] >> ((r - 1) & 63)) & 1;   // coeff of x^(r-1)
    if (w == 0) v ^= wrap;
    if (w == (size_t)(s >> 6)) v ^= wrap << (s & 63);
    dst[w] ^= v;
}

__global__ void k_set_word0(u64 *p, size_t nw, u64 v0) {
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nw) p[w] = w ? 0 : v0;
}

__global__ void k_xor_word0(u64 *p, u64 v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) p[0] ^= v;
}

// r below 2^34: the kernels address a residue with 32-bit byte offsets
bool valid_rs(u64 r, u64 s) { return r >= 3 && r < (1ULL << 34) && s >= 1 && 2 * s < r; }

// tuning knobs of the batched squaring (gf2t_tune / GF2T_SQ_PPT, GF2T_SQ_ORDER)
int g_sq_ppt = -1, g_sq_order = -1, g_sq_nt = -1, g_sq_fuse = -1;
// two squarings per launch wherever a chain has two steps left
bool sq_fuse() {
    if (g_sq_fuse < 0) {
        const char *e = getenv("GF2T_SQ_FUSE");
        g_sq_fuse = e ? atoi(e) : 1;
    }
    return g_sq_fuse != 0;
}
bool sq_nt() {
    if (g_sq_nt < 0) {
        const char *e = getenv("GF2T_SQ_NT");
        g_sq_nt = e ? atoi(e) : 1;
    }
    return g_sq_nt != 0;
}
void sq_tuning(int &ppt, int &order) {
    if (g_sq_ppt < 0) {
        const char *e = getenv("GF2T_SQ_PPT");
        g_sq_ppt = e ? atoi(e) : 2;
    }
    if (g_sq_order < 0) {
        const char *e = getenv("GF2T_SQ_ORDER");
        g_sq_order = e ? atoi(e) : 0;
    }
    ppt = (g_sq_ppt == 2 || g_sq_ppt == 4) ? g_sq_ppt : 1;
    order = g_sq_order ? 1 : 0;
}

// 16-byte vector copy, four loads in flight per lane
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_copy(const ulonglong2 *__restrict__ src,
                                                     ulonglong2 *__restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        ulonglong2 v[4] = {src[i], src[i + stride], src[i + 2 * stride], src[i + 3 * stride]};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (NT) {
                typedef u64 v2u __attribute__((ext_vector_type(2)));
                __builtin_nontemporal_store(v2u{v[k].x, v[k].y}, (v2u *)&dst[i + k * stride]);
            } else {
                dst[i + k * stride] = v[k];
            }
        }
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

unsigned blocks_for(size_t work, int threads) {
    return (unsigned)((work + threads - 1) / threads);
}

}  // namespace

extern "C" int gf2t_square_fold(const uint64_t *din, uint64_t *dout, uint64_t r,
                                uint64_t s, void *stream) {
    if (!valid_rs(r, s) || !din || !dout || din == dout) {
        gf2t::set_error("gf2t_square_fold: need r >= 3, 1 <= s, 2s < r, din != dout");
        return GF2T_EINVAL;
    }
    SqConst c = sq_const(r, s);
    if (sq_w12_ok(r))
        hipLaunchKernelGGL(k_square_fold<true>, dim3(blocks_for((c.nw + 1) / 2, SQ_THREADS)),
                           dim3(SQ_THREADS), 0, (hipStream_t)stream, din, dout, c);
    else
        hipLaunchKernelGGL(k_square_fold<false>, dim3(blocks_for((c.nw + 1) / 2, SQ_THREADS)),
                           dim3(SQ_THREADS), 0, (hipStream_t)stream, din, dout, c);
    GF2T_CHECK_LAUNCH("k_square_fold");
    return GF2T_OK;
}

extern "C" int gf2t_square_fold_steps(uint64_t *dres, uint64_t *dtmp, uint64_t r,
                                      uint64_t s, uint64_t steps, void *stream) {
    if (!valid_rs(r, s) || !dres || !dtmp || dres == dtmp) {
        gf2t::set_error("gf2t_square_fold_steps: bad arguments");
        return GF2T_EINVAL;
    }
    SqConst c = sq_const(r, s);
    size_t nw = c.nw;
    hipStream_t st = (hipStream_t)stream;
    u64 *a = dres, *b = dtmp;
    const bool fuse = sq_fuse(), w12 = sq_w12_ok(r);
    const Sq2Chain h2 = sq2_chain(c);
    const dim3 g2((unsigned)sq2_units(h2)), g1(blocks_for((nw + 1) / 2, SQ_THREADS));
    u64 launches = 0;
    for (u64 k = 0; k < steps; launches++) {
        if (fuse && steps - k >= 2) {
            if (w12)
                hipLaunchKernelGGL(k_square2_fold<true>, g2, dim3(SQ_THREADS), 0, st, a, b, c, h2);
            else
                hipLaunchKernelGGL(k_square2_fold<false>, g2, dim3(SQ_THREADS), 0, st, a, b, c, h2);
            k += 2;
        } else {
            if (w12)
                hipLaunchKernelGGL(k_square_fold<true>, g1, dim3(SQ_THREADS), 0, st, a, b, c);
            else
                hipLaunchKernelGGL(k_square_fold<false>, g1, dim3(SQ_THREADS), 0, st, a, b, c);
            k += 1;
        }
        std::swap(a, b);
    }
    GF2T_CHECK_LAUNCH("k_square_fold");
    if (launches & 1) GF2T_HIP(hipMemcpyAsync(dres, dtmp, nw * 8, hipMemcpyDeviceToDevice, st));
    return GF2T_OK;
}

namespace {
// descriptor and tile-order tables for the batched chain, cached and
// grown on demand
struct BatchScratch {
    SqDesc *d = nullptr;
    int *ord = nullptr;
    size_t cap = 0, ocap = 0;
    int dev = -1;
} g_batch;

template <int PPT, bool W12>
void launch_batch(unsigned gx, unsigned active, const SqDesc *tab, hipStream_t st) {
    if (sq_nt())
        hipLaunchKernelGGL((k_square_fold_batch<PPT, true, W12>), dim3(gx, active),
                           dim3(SQ_THREADS), 0, st, tab);
    else
        hipLaunchKernelGGL((k_square_fold_batch<PPT, false, W12>), dim3(gx, active),
                           dim3(SQ_THREADS), 0, st, tab);
}
template <bool W12>
void launch_batch2(unsigned gx, unsigned active, const SqDesc *tab, hipStream_t st) {
    if (sq_nt())
        hipLaunchKernelGGL((k_square2_fold_batch<true, W12>), dim3(gx, active), dim3(SQ_THREADS),
                           0, st, tab);
    else
        hipLaunchKernelGGL((k_square2_fold_batch<false, W12>), dim3(gx, active),
                           dim3(SQ_THREADS), 0, st, tab);
}
}  // namespace

namespace gf2t_host {
bool square_fuse_enabled() { return sq_fuse(); }
}

int gf2t_fft_set_tile_bits(int m, int64_t T);   // fft.hip
int gf2t_fft_tile_bits(int m);
int gf2t_fft_set_ept(int64_t ept);                // fft.hip
int gf2t_fft_ept();
int gf2t_fft_set_low_bits(int64_t c);             // fft.hip
int gf2t_fft_low_bits();
int gf2t_hm_set_zc_words(int64_t words);          // fft.hip
int64_t gf2t_hm_zc_words();
int gf2t_sieve_tune(const char *key, int64_t v);   // sieve.hip
int gf2t_sieve_tune_get(const char *key, int64_t *v);

extern "C" int gf2t_tune(const char *key, int64_t value) {
    if (!key) return GF2T_EINVAL;
    if (!strncmp(key, "fft_tile_", 9) && key[9]) {
        char *end = nullptr;
        long m = strtol(key + 9, &end, 10);
        if (*end == 0) return gf2t_fft_set_tile_bits((int)m, value);
    }
    if (!strcmp(key, "sieve_wide_from") || !strcmp(key, "sieve_slices"))
        return gf2t_sieve_tune(key, value);
    if (!strncmp(key, "block_m_", 8) && key[8]) {
        char *end = nullptr;
        long k = strtol(key + 8, &end, 10);
        if (*end == 0 && gf2t_host::set_tuned_block_m((int)k, (int)value)) return GF2T_OK;
        gf2t::set_error("gf2t_tune: block_m_<k> needs 2 <= k <= 40 and 0 <= m <= 64");
        return GF2T_EINVAL;
    }
    if (!strcmp(key, "sq_ppt")) { g_sq_ppt = (int)value; return GF2T_OK; }
    if (!strcmp(key, "sq_order")) { g_sq_order = (int)value; return GF2T_OK; }
    if (!strcmp(key, "sq_nt")) { g_sq_nt = (int)value; return GF2T_OK; }
    if (!strcmp(key, "sq_fuse")) { g_sq_fuse = (int)value; return GF2T_OK; }
    if (!strcmp(key, "gcd_dev_bits")) {
        gf2t_host::set_device_product_bits(value > 0 ? (uint64_t)value : 0);
        return GF2T_OK;
    }
    if (!strcmp(key, "gcd_base_bits")) {
        if (gf2t_host::set_gcd_base_bits(value)) return GF2T_OK;
        gf2t::set_error("gf2t_tune: gcd_base_bits needs 128 <= bits <= 65536");
        return GF2T_EINVAL;
    }
    if (!strcmp(key, "gcd_school_words")) {
        if (gf2t_host::set_school_words(value)) return GF2T_OK;
        gf2t::set_error("gf2t_tune: gcd_school_words needs 8 <= words <= 512");
        return GF2T_EINVAL;
    }
    if (!strcmp(key, "fft_ept")) return gf2t_fft_set_ept(value);
    if (!strcmp(key, "fft_low_bits")) return gf2t_fft_set_low_bits(value);
    if (!strcmp(key, "hm_zc_words")) return gf2t_hm_set_zc_words(value);
    if (!strcmp(key, "scan_stagger")) {
        if (value < -1 || value > 64) {
            gf2t::set_error("gf2t_tune: scan_stagger needs -1 <= value <= 64");
            return GF2T_EINVAL;
        }
        gf2t_host::set_scan_stagger((int)value);
        return GF2T_OK;
    }
    gf2t::set_error("gf2t_tune: unknown key '%s'", key);
    return GF2T_EINVAL;
}

extern "C" int gf2t_tune_get(const char *key, int64_t *value) {
    if (!key || !value) return GF2T_EINVAL;
    int ppt = 0, order = 0;
    sq_tuning(ppt, order);
    if (!strcmp(key, "sq_ppt")) { *value = ppt; return GF2T_OK; }
    if (!strcmp(key, "sq_order")) { *value = order; return GF2T_OK; }
    if (!strcmp(key, "sq_nt")) { *value = sq_nt() ? 1 : 0; return GF2T_OK; }
    if (!strcmp(key, "sq_fuse")) { *value = sq_fuse() ? 1 : 0; return GF2T_OK; }
    if (!strcmp(key, "gcd_dev_bits")) {
        *value = (int64_t)gf2t_host::device_product_bits();
        return GF2T_OK;
    }
    if (!strcmp(key, "gcd_base_bits")) {
        *value = gf2t_host::gcd_base_bits();
        return GF2T_OK;
    }
    if (!strcmp(key, "gcd_school_words")) {
        *value = gf2t_host::school_words();
        return GF2T_OK;
    }
    if (!strcmp(key, "fft_ept")) {
        *value = gf2t_fft_ept();
        return GF2T_OK;
    }
    if (!strcmp(key, "fft_low_bits")) {
        *value = gf2t_fft_low_bits();
        return GF2T_OK;
    }
    if (!strcmp(key, "hm_zc_words")) {
        *value = gf2t_hm_zc_words();
        return GF2T_OK;
    }
    if (!strcmp(key, "scan_stagger")) {
        *value = gf2t_host::scan_stagger();
        return GF2T_OK;
    }
    if (!strcmp(key, "sieve_wide_from") || !strcmp(key, "sieve_slices"))
        return gf2t_sieve_tune_get(key, value);
    if (!strncmp(key, "fft_tile_", 9) && key[9]) {
        char *end = nullptr;
        long m = strtol(key + 9, &end, 10);
        if (*end == 0 && gf2t_fft_tile_bits((int)m)) {
            *value = gf2t_fft_tile_bits((int)m);
            return GF2T_OK;
        }
    }
    if (!strncmp(key, "block_m_", 8) && key[8]) {
        char *end = nullptr;
        long k = strtol(key + 8, &end, 10);
        if (*end == 0 && k >= 2 && k <= 40) {
            *value = gf2t_host::tuned_block_m(1ULL << k);
            return GF2T_OK;
        }
    }
    gf2t::set_error("gf2t_tune_get: unknown key '%s'", key);
    return GF2T_EINVAL;
}

extern "C" int gf2t_square_fold_batch(uint64_t *const *dres, uint64_t *const *dtmp,
                                      const uint64_t *r, const uint64_t *s,
                                      const uint64_t *steps, int n, void *stream) {
    if (n < 0 || (n > 0 && (!dres || !dtmp || !r || !s || !steps))) {
        gf2t::set_error("gf2t_square_fold_batch: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    if (n > 65535) {
        gf2t::set_error("gf2t_square_fold_batch: at most 65535 residues per call");
        return GF2T_ERANGE;
    }
    for (int i = 0; i < n; i++)
        if (!valid_rs(r[i], s[i]) || !dres[i] || !dtmp[i] || dres[i] == dtmp[i] ||
            (((uintptr_t)dres[i] | (uintptr_t)dtmp[i]) & 15)) {
            gf2t::set_error("gf2t_square_fold_batch: residue %d has bad r/s/buffers "
                            "(distinct, 16-byte aligned)", i);
            return GF2T_EINVAL;
        }
    hipStream_t st = (hipStream_t)stream;
    int ppt, order;
    sq_tuning(ppt, order);
    // order residues by remaining steps, longest first: the residues
    // still active at step k are a prefix, so each launch covers exactly
    // the live chains (uneven frontiers, per-s early exit)
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return steps[a] > steps[b]; });
    int dev = 0;
    GF2T_HIP(hipGetDevice(&dev));
    size_t need = 2 * (size_t)n;
    // per-residue tile orders
    std::vector<int> otab;
    std::vector<size_t> ooff(n, 0);
    size_t gx = 0;
    const u64 tile_pairs = (u64)SQ_THREADS * ppt;
    for (int k = 0; k < n; k++) {
        int i = ord[k];
        SqConst c = sq_const(r[i], s[i]);
        u64 T = ((c.nw + 1) / 2 + tile_pairs - 1) / tile_pairs;
        if (order) {
            TileOrder o = tile_order(T, sq_delta_pairs(c), tile_pairs);
            ooff[k] = otab.size();
            for (u64 x = 0; x < 8 * o.P; x++) otab.push_back((int)tile_of(o, x));
            gx = std::max<size_t>(gx, 8 * o.P);
        } else {
            gx = std::max<size_t>(gx, T);
        }
    }
    if (g_batch.dev != dev) {
        g_batch = BatchScratch();
        g_batch.dev = dev;
    }
    if (g_batch.cap < need) {
        if (g_batch.d) (void)hipFree(g_batch.d);
        g_batch.d = nullptr;
        GF2T_HIP(hipMalloc(&g_batch.d, need * sizeof(SqDesc)));
        g_batch.cap = need;
    }
    if (order && g_batch.ocap < otab.size()) {
        if (g_batch.ord) (void)hipFree(g_batch.ord);
        g_batch.ord = nullptr;
        GF2T_HIP(hipMalloc(&g_batch.ord, otab.size() * sizeof(int)));
        g_batch.ocap = otab.size();
    }
    // a block index beyond a residue's own table must idle: pad each
    // residue's table to gx entries
    if (order) {
        std::vector<int> padded((size_t)n * gx, -1);
        for (int k = 0; k < n; k++) {
            size_t len = (k + 1 < n ? ooff[k + 1] : otab.size()) - ooff[k];
            std::copy(otab.begin() + ooff[k], otab.begin() + ooff[k] + len,
                      padded.begin() + (size_t)k * gx);
        }
        otab.swap(padded);
        if (g_batch.ocap < otab.size()) {
            if (g_batch.ord) (void)hipFree(g_batch.ord);
            g_batch.ord = nullptr;
            GF2T_HIP(hipMalloc(&g_batch.ord, otab.size() * sizeof(int)));
            g_batch.ocap = otab.size();
        }
        GF2T_HIP(hipMemcpyAsync(g_batch.ord, otab.data(), otab.size() * sizeof(int),
                                hipMemcpyHostToDevice, st));
    }
    std::vector<SqDesc> h(need);
    for (int k = 0; k < n; k++) {
        int i = ord[k];
        SqConst c = sq_const(r[i], s[i]);
        const int *ot = order ? g_batch.ord + (size_t)k * gx : nullptr;
        const Sq2Chain h2 = sq2_chain(c);
        h[k] = SqDesc{dres[i], dtmp[i], c, ot, h2};          // even steps
        h[n + k] = SqDesc{dtmp[i], dres[i], c, ot, h2};      // odd steps
    }
    GF2T_HIP(hipMemcpyAsync(g_batch.d, h.data(), need * sizeof(SqDesc), hipMemcpyHostToDevice, st));
    // Every launch advances all live chains together, by two squarings
    // while each of them has two left, else by one; a chain leaves the
    // prefix when its steps are done, after launches[k] launches (its
    // value is in dtmp when that count is odd).
    const bool fuse = sq_fuse();
    u64 gx2 = 0;   // work units of the residue with the most
    bool w12 = true;   // 12-byte window loads: no r of the batch a multiple of 64
    for (int i = 0; i < n; i++) {
        gx2 = std::max(gx2, sq2_units(sq2_chain(sq_const(r[i], s[i]))));
        w12 = w12 && sq_w12_ok(r[i]);
    }
    std::vector<u64> launches(n, 0);
    int active = n;
    u64 nl = 0;
    for (u64 k = 0;; nl++) {
        while (active > 0 && steps[ord[active - 1]] <= k) launches[--active] = nl;
        if (!active) break;
        const SqDesc *tab = g_batch.d + ((nl & 1) ? n : 0);
        if (fuse && steps[ord[active - 1]] - k >= 2) {
            if (w12) launch_batch2<true>((unsigned)gx2, active, tab, st);
            else launch_batch2<false>((unsigned)gx2, active, tab, st);
            k += 2;
        } else {
            if (ppt == 4 && w12) launch_batch<4, true>((unsigned)gx, active, tab, st);
            else if (ppt == 4) launch_batch<4, false>((unsigned)gx, active, tab, st);
            else if (ppt == 2 && w12) launch_batch<2, true>((unsigned)gx, active, tab, st);
            else if (ppt == 2) launch_batch<2, false>((unsigned)gx, active, tab, st);
            else if (w12) launch_batch<1, true>((unsigned)gx, active, tab, st);
            else launch_batch<1, false>((unsigned)gx, active, tab, st);
            k += 1;
        }
    }
    GF2T_CHECK_LAUNCH("k_square_fold_batch");
    for (int k = 0; k < n; k++)
        if (launches[k] & 1)
            GF2T_HIP(hipMemcpyAsync(dres[ord[k]], dtmp[ord[k]], words_of(r[ord[k]]) * 8,
                                    hipMemcpyDeviceToDevice, st));
    // the host tables must outlive the async copies
    GF2T_HIP(hipStreamSynchronize(st));
    return GF2T_OK;
}

extern "C" int gf2t_fold(const uint64_t *dsrc, uint64_t src_bits, uint64_t *dout,
                         uint64_t r, uint64_t s, void *stream) {
    if (!valid_rs(r, s) || !dsrc || !dout || dsrc == dout || src_bits > 2 * r - 1) {
        gf2t::set_error("gf2t_fold: need valid r/s, src_bits <= 2r-1, dsrc != dout");
        return GF2T_EINVAL;
    }
    size_t nw = words_of(r), pw = words_of(src_bits);
    hipLaunchKernelGGL(k_fold, dim3(blocks_for(nw, SQ_THREADS)), dim3(SQ_THREADS), 0,
                       (hipStream_t)stream, dsrc, pw, src_bits, dout, nw, r, s);
    GF2T_CHECK_LAUNCH("k_fold");
    return GF2T_OK;
}

extern "C" int gf2t_mulx_xor(uint64_t *ddst, const uint64_t *dsrc, uint64_t r,
                             uint64_t s, void *stream) {
    if (!valid_rs(r, s) || !ddst || !dsrc || ddst == dsrc) {
        gf2t::set_error("gf2t_mulx_xor: bad arguments");
        return GF2T_EINVAL;
    }
    size_t nw = words_of(r);
    hipLaunchKernelGGL(k_mulx_xor, dim3(blocks_for(nw, SQ_THREADS)), dim3(SQ_THREADS), 0,
                       (hipStream_t)stream, ddst, dsrc, nw, r, s);
    GF2T_CHECK_LAUNCH("k_mulx_xor");
    return GF2T_OK;
}

extern "C" int gf2t_set_x(uint64_t *dp, size_t nw, void *stream) {
    if (!dp || nw == 0) { gf2t::set_error("gf2t_set_x: bad arguments"); return GF2T_EINVAL; }
    hipLaunchKernelGGL(k_set_word0, dim3(blocks_for(nw, 256)), dim3(256), 0,
                       (hipStream_t)stream, dp, nw, 2ULL);
    GF2T_CHECK_LAUNCH("k_set_word0");
    return GF2T_OK;
}

extern "C" int gf2t_set_one(uint64_t *dp, size_t nw, void *stream) {
    if (!dp || nw == 0) { gf2t::set_error("gf2t_set_one: bad arguments"); return GF2T_EINVAL; }
    hipLaunchKernelGGL(k_set_word0, dim3(blocks_for(nw, 256)), dim3(256), 0,
                       (hipStream_t)stream, dp, nw, 1ULL);
    GF2T_CHECK_LAUNCH("k_set_word0");
    return GF2T_OK;
}

extern "C" int gf2t_xor_x(uint64_t *dp, void *stream) {
    if (!dp) { gf2t::set_error("gf2t_xor_x: null pointer"); return GF2T_EINVAL; }
    hipLaunchKernelGGL(k_xor_word0, dim3(1), dim3(64), 0, (hipStream_t)stream, dp, 2ULL);
    GF2T_CHECK_LAUNCH("k_xor_word0");
    return GF2T_OK;
}

extern "C" int gf2t_square_fold_slab(const uint64_t *din, uint64_t *dout, size_t pitch,
                                     uint64_t r, const uint64_t *ds, int n, void *stream) {
    size_t nw = words_of(r);
    if (!din || !dout || !ds || din == dout || r < 3 || pitch < nw || n < 0 || n > 65535) {
        gf2t::set_error("gf2t_square_fold_slab: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    unsigned gx = blocks_for((nw + 1) / 2, SQ_THREADS);
    hipLaunchKernelGGL(k_square_fold_slab, dim3(gx, n), dim3(SQ_THREADS), 0,
                       (hipStream_t)stream, din, dout, (u64)pitch, r, ds);
    GF2T_CHECK_LAUNCH("k_square_fold_slab");
    return GF2T_OK;
}

extern "C" int gf2t_stream_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if (!dst || !src || (bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15)) {
        gf2t::set_error("gf2t_stream_copy: 16-byte aligned pointers and size required");
        return GF2T_EINVAL;
    }
    size_t n = bytes / 16;
    // one 4-load iteration per lane: the widest grid reads fastest here
    // (5.46 TB/s copying 4.36 GB, against 4.87 TB/s with 8192 blocks)
    unsigned blocks = (unsigned)std::min<size_t>((n + 1023) / 1024, (size_t)1 << 20);
    if (sq_nt())
        hipLaunchKernelGGL(k_stream_copy<true>, dim3(blocks ? blocks : 1), dim3(256), 0,
                           (hipStream_t)stream, (const ulonglong2 *)src, (ulonglong2 *)dst, n);
    else
        hipLaunchKernelGGL(k_stream_copy<false>, dim3(blocks ? blocks : 1), dim3(256), 0,
                           (hipStream_t)stream, (const ulonglong2 *)src, (ulonglong2 *)dst, n);
    GF2T_CHECK_LAUNCH("k_stream_copy");
    return GF2T_OK;
}

extern "C" int gf2t_square2_fold_slab_g(const uint64_t *din, uint64_t *dout, size_t pitch,
                                        uint64_t r, const uint64_t *ds, size_t n, int group,
                                        void *stream) {
    size_t nw = words_of(r);
    if (!din || !dout || !ds || din == dout || r < 5 || r >= (1ULL << 34) || pitch < nw ||
        (pitch & 1) || (((uintptr_t)din | (uintptr_t)dout) & 15) || group < 1 ||
        n > 0xFFFFFFFFu) {
        gf2t::set_error("gf2t_square2_fold_slab_g: bad arguments (even pitch, 16-byte "
                        "aligned slabs)");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    const unsigned gx = (unsigned)(((nw + 1) / 2 + SQ2_TP - 1) / SQ2_TP);
    const unsigned gy = n < 65535 ? (unsigned)n : 65535u;
    const unsigned gz = (unsigned)((n + gy - 1) / gy);
    if (sq_w12_ok(r))
        hipLaunchKernelGGL(k_square2_fold_slab_g<true>, dim3(gx, gy, gz), dim3(SQ_THREADS), 0,
                           (hipStream_t)stream, din, dout, (u64)pitch, r, ds, (unsigned)n,
                           (unsigned)group);
    else
        hipLaunchKernelGGL(k_square2_fold_slab_g<false>, dim3(gx, gy, gz), dim3(SQ_THREADS), 0,
                           (hipStream_t)stream, din, dout, (u64)pitch, r, ds, (unsigned)n,
                           (unsigned)group);
    GF2T_CHECK_LAUNCH("k_square2_fold_slab_g");
    return GF2T_OK;
}

// ---- two squarings of a slab in work units, with per-group layouts
struct gf2t_sq2_slab_plan {
    u64 r = 0;
    size_t cap = 0, ngroups = 0;
    unsigned gx = 0;   // work units of the group with the most
    Sq2SlabDesc *d = nullptr;
    std::vector<Sq2SlabDesc> h;   // host copy (source of the last upload)
};

extern "C" int gf2t_sq2_slab_plan_set(gf2t_sq2_slab_plan *p, const uint64_t *hs, size_t ngroups,
                                      void *stream) {
    if (!p || !hs || ngroups == 0 || ngroups > p->cap) {
        gf2t::set_error("gf2t_sq2_slab_plan_set: bad arguments (at most %zu groups)",
                        p ? p->cap : (size_t)0);
        return GF2T_EINVAL;
    }
    for (size_t g = 0; g < ngroups; g++)
        if (!valid_rs(p->r, hs[g])) {
            gf2t::set_error("gf2t_sq2_slab_plan_set: group %zu has bad s", g);
            return GF2T_EINVAL;
        }
    // the previous upload may still be in flight from h: let it finish
    // before h changes
    GF2T_HIP(hipStreamSynchronize((hipStream_t)stream));
    unsigned gx = 0;
    for (size_t g = 0; g < ngroups; g++) {
        const SqConst c = sq_const(p->r, hs[g]);
        p->h[g] = Sq2SlabDesc{c, sq2_chain(c)};
        gx = std::max(gx, (unsigned)sq2_units(p->h[g].h));
    }
    GF2T_HIP(hipMemcpyAsync(p->d, p->h.data(), ngroups * sizeof(Sq2SlabDesc),
                            hipMemcpyHostToDevice, (hipStream_t)stream));
    p->ngroups = ngroups;
    p->gx = gx;
    return GF2T_OK;
}

extern "C" int gf2t_sq2_slab_plan_create(uint64_t r, const uint64_t *hs, size_t ngroups,
                                         gf2t_sq2_slab_plan **plan) {
    if (!plan || !hs || ngroups == 0 || r < 5 || r >= (1ULL << 34)) {
        gf2t::set_error("gf2t_sq2_slab_plan_create: bad arguments");
        return GF2T_EINVAL;
    }
    *plan = nullptr;
    gf2t_sq2_slab_plan *p = new gf2t_sq2_slab_plan;
    p->r = r;
    p->cap = ngroups;
    p->h.resize(ngroups);
    hipError_t e = hipMalloc((void **)&p->d, ngroups * sizeof(Sq2SlabDesc));
    if (e != hipSuccess) {
        delete p;
        return gf2t::hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    int rc = gf2t_sq2_slab_plan_set(p, hs, ngroups, nullptr);
    if (rc == GF2T_OK) rc = hipStreamSynchronize(nullptr) == hipSuccess ? GF2T_OK : GF2T_EHIP;
    if (rc) {
        (void)hipFree(p->d);
        delete p;
        return rc;
    }
    *plan = p;
    return GF2T_OK;
}

extern "C" int gf2t_sq2_slab_plan_destroy(gf2t_sq2_slab_plan *p) {
    if (!p) return GF2T_OK;
    if (p->d) (void)hipFree(p->d);
    delete p;
    return GF2T_OK;
}

extern "C" int gf2t_square2_fold_slab_p(const gf2t_sq2_slab_plan *p, const uint64_t *din,
                                        uint64_t *dout, size_t pitch, size_t n, int group,
                                        void *stream) {
    if (!p || !din || !dout || din == dout || group < 1 || pitch < words_of(p->r) ||
        (pitch & 1) || (((uintptr_t)din | (uintptr_t)dout) & 15) || n > 0xFFFFFFFFu ||
        (n + (size_t)group - 1) / (size_t)group > p->ngroups) {
        gf2t::set_error("gf2t_square2_fold_slab_p: bad arguments (even pitch, 16-byte aligned "
                        "slabs, n / group within the plan's groups)");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    const unsigned gy = n < 65535 ? (unsigned)n : 65535u;
    const unsigned gz = (unsigned)((n + gy - 1) / gy);
    const dim3 grid(p->gx, gy, gz);
    const bool w12 = sq_w12_ok(p->r), nt = sq_nt();
    hipStream_t st = (hipStream_t)stream;
    if (nt && w12)
        hipLaunchKernelGGL((k_square2_fold_slab_u<true, true>), grid, dim3(SQ_THREADS), 0, st, din,
                           dout, (u64)pitch, p->d, (unsigned)n, (unsigned)group);
    else if (nt)
        hipLaunchKernelGGL((k_square2_fold_slab_u<true, false>), grid, dim3(SQ_THREADS), 0, st,
                           din, dout, (u64)pitch, p->d, (unsigned)n, (unsigned)group);
    else if (w12)
        hipLaunchKernelGGL((k_square2_fold_slab_u<false, true>), grid, dim3(SQ_THREADS), 0, st,
                           din, dout, (u64)pitch, p->d, (unsigned)n, (unsigned)group);
    else
        hipLaunchKernelGGL((k_square2_fold_slab_u<false, false>), grid, dim3(SQ_THREADS), 0, st,
                           din, dout, (u64)pitch, p->d, (unsigned)n, (unsigned)group);
    GF2T_CHECK_LAUNCH("k_square2_fold_slab_u");
    return GF2T_OK;
}
