// ddf.hip -- blocked distinct-degree factorization on slabs of residues
// (the h-table scheme of the reference's tsfactor, src/tsfactor.cu:71-126,
// :345-365), for a cohort of s values at once.
//
// For a block of m consecutive degrees k+1..k+m the product
//     C = prod_{t<m} (x + x^(2^(k+1+t)))  =  x^m + sum_j h_j x^(m-1-j)
// has coefficients h_j = e_(j+1)(x^(2^(k+1)), ..., x^(2^(k+m))), the
// elementary symmetric functions of the m Frobenius iterates.  Frobenius
// is a ring map, so advancing the block by one degree squares every h_j:
// per degree m squarings, and per block one Horner pass plus one product
// ACC <- ACC * C instead of m products.
//
// h-table layout: slot j of residue i at H + i*res_pitch + j*slot_pitch
// (res_pitch = m*slot_pitch), so the whole table is a slab of n*m
// residues with pitch slot_pitch whose residue y uses s_(y / m).
#include "gf2t_internal.hpp"
#include "host_poly.hpp"
#include "../../include/gf2t.h"
#include <vector>

namespace {

constexpr int DDF_THREADS = 256;
constexpr int DDF_MAX_M = 64;

unsigned blocks_for(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

// squaring of a slab in groups: residue y (grid-stride over y) uses
// svec[y / group]; W12 as in the squaring kernels (12-byte window loads,
// for r not a multiple of 64)
template <bool W12>
__global__ __launch_bounds__(DDF_THREADS) void k_square_fold_slab_g(
        const u64 *__restrict__ in, u64 *__restrict__ out, u64 pitch, u64 r,
        const u64 *__restrict__ svec, unsigned n, unsigned group) {
    for (unsigned y = blockIdx.y; y < n; y += gridDim.y) {
        const SqConst c = sq_const(r, svec[y / group]);
        size_t i = (size_t)blockIdx.x * DDF_THREADS + threadIdx.x;
        if (i >= (c.nw + 1) / 2) continue;
        const u64 *src = in + (u64)y * pitch;
        u64 *dst = out + (u64)y * pitch;
        u64 o0, o1;
        square_fold_pair_fast<W12>(src, c, i, o0, o1);
        dst[2 * i] = o0;
        if (2 * i + 1 < c.nw) dst[2 * i + 1] = o1;
    }
}

// dst_y ^= x * src_y mod x^r + x^(s_y) + 1
__global__ __launch_bounds__(DDF_THREADS) void k_mulx_xor_slab(
        u64 *__restrict__ dst, u64 dpitch, const u64 *__restrict__ src, u64 spitch, size_t nw,
        u64 r, const u64 *__restrict__ svec, unsigned n) {
    size_t w = (size_t)blockIdx.x * DDF_THREADS + threadIdx.x;
    if (w >= nw) return;
    for (unsigned y = blockIdx.y; y < n; y += gridDim.y) {
        const u64 *a = src + (u64)y * spitch;
        const u64 s = svec[y];
        u64 v = (a[w] << 1) | (w ? a[w - 1] >> 63 : 0);
        if (w == (size_t)(r >> 6)) v &= ~(1ULL << (r & 63));
        u64 wrap = (a[(r - 1) >> 6] >> ((r - 1) & 63)) & 1;   // coefficient of x^(r-1)
        if (w == 0) v ^= wrap;
        if (w == (size_t)(s >> 6)) v ^= wrap << (s & 63);
        dst[(u64)y * dpitch + w] ^= v;
    }
}

__global__ void k_xor_word0_slab(u64 *p, u64 pitch, u64 v, unsigned n) {
    unsigned y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y < n) p[(u64)y * pitch] ^= v;
}

__global__ __launch_bounds__(DDF_THREADS) void k_set_one_slab(u64 *__restrict__ p, u64 pitch,
                                                              size_t nw, unsigned n) {
    size_t w = (size_t)blockIdx.x * DDF_THREADS + threadIdx.x;
    if (w >= nw) return;
    for (unsigned y = blockIdx.y; y < n; y += gridDim.y) p[(u64)y * pitch + w] = w ? 0 : 1;
}

__global__ __launch_bounds__(DDF_THREADS) void k_copy_slab(
        u64 *__restrict__ dst, u64 dpitch, const u64 *__restrict__ src, u64 spitch, size_t nw,
        unsigned n) {
    size_t w = (size_t)blockIdx.x * DDF_THREADS + threadIdx.x;
    if (w >= nw) return;
    for (unsigned y = blockIdx.y; y < n; y += gridDim.y)
        dst[(u64)y * dpitch + w] = src[(u64)y * spitch + w];
}

// bits [p, p + len) of a residue (len <= 63, p + len <= r)
__device__ inline u64 bits_at(const u64 *h, size_t nw, u64 p, unsigned len) {
    size_t wi = (size_t)(p >> 6);
    unsigned off = (unsigned)(p & 63);
    u64 v = h[wi] >> off;
    if (off && wi + 1 < nw) v |= h[wi + 1] << (64 - off);
    return len ? (v & ((1ULL << len) - 1)) : 0;
}

// C_y = x^m + sum_j h_(y,j) x^(m-1-j) mod x^r + x^(s_y) + 1 in gather form.
// The unreduced sum has bits up to r+m-2; its excess E = bits [r, r+m-2]
// folds once to E + E x^s (x^(r+e) = x^(s+e) + x^e), which stays below r.
__global__ __launch_bounds__(DDF_THREADS) void k_horner_slab(
        const u64 *__restrict__ H, u64 slot_pitch, u64 res_pitch, int m, size_t nw, u64 r,
        const u64 *__restrict__ svec, u64 *__restrict__ C, u64 cpitch, unsigned n) {
    __shared__ u64 ex[64];
    const size_t w = (size_t)blockIdx.x * DDF_THREADS + threadIdx.x;
    const unsigned top = (unsigned)(r & 63);
    const size_t topw = (size_t)(r >> 6);
    for (unsigned y = blockIdx.y; y < n; y += gridDim.y) {
        const u64 *h = H + (u64)y * res_pitch;
        const u64 s = svec[y];
        // excess: (h_j x^sh) >> r = the top sh bits of h_j, sh = m-1-j
        if (threadIdx.x < (unsigned)m) {
            unsigned sh = (unsigned)(m - 1 - (int)threadIdx.x);
            ex[threadIdx.x] = sh ? bits_at(h + (u64)threadIdx.x * slot_pitch, nw, r - sh, sh) : 0;
        }
        __syncthreads();
        u64 E = 0;
        for (int j = 0; j < m; j++) E ^= ex[j];
        if (w < nw) {
            u64 acc = 0;
            for (int j = 0; j < m; j++) {
                const u64 *hj = h + (u64)j * slot_pitch;
                unsigned sh = (unsigned)(m - 1 - j);
                u64 a = hj[w];
                u64 b = w ? hj[w - 1] : 0;
                acc ^= sh ? ((a << sh) | (b >> (64 - sh))) : a;
            }
            if (w == topw) acc &= (1ULL << top) - 1;   // w == topw only when top != 0
            if (w == 0) acc ^= E;                      // E has m - 1 <= 63 bits
            if (w == (size_t)(m >> 6)) acc ^= 1ULL << (m & 63);   // x^m (2m < r)
            const size_t sw = (size_t)(s >> 6);
            const unsigned so = (unsigned)(s & 63);
            if (w == sw) acc ^= E << so;
            if (so && w == sw + 1) acc ^= E >> (64 - so);
            C[(u64)y * cpitch + w] = acc;
        }
        __syncthreads();
    }
}

unsigned grid_y(unsigned n) { return n < 65535u ? n : 65535u; }

}  // namespace

extern "C" int gf2t_square_fold_slab_g(const uint64_t *din, uint64_t *dout, size_t pitch,
                                       uint64_t r, const uint64_t *ds, size_t n, int group,
                                       void *stream) {
    size_t nw = words_of(r);
    if (!din || !dout || !ds || din == dout || r < 3 || pitch < nw || group < 1 ||
        n > 0xFFFFFFFFu) {
        gf2t::set_error("gf2t_square_fold_slab_g: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    unsigned gx = blocks_for((nw + 1) / 2, DDF_THREADS);
    if (sq_w12_ok(r))
        hipLaunchKernelGGL(k_square_fold_slab_g<true>, dim3(gx, grid_y((unsigned)n)),
                           dim3(DDF_THREADS), 0, (hipStream_t)stream, din, dout, (u64)pitch, r,
                           ds, (unsigned)n, (unsigned)group);
    else
        hipLaunchKernelGGL(k_square_fold_slab_g<false>, dim3(gx, grid_y((unsigned)n)),
                           dim3(DDF_THREADS), 0, (hipStream_t)stream, din, dout, (u64)pitch, r,
                           ds, (unsigned)n, (unsigned)group);
    GF2T_CHECK_LAUNCH("k_square_fold_slab_g");
    return GF2T_OK;
}

extern "C" int gf2t_mulx_xor_slab(uint64_t *ddst, size_t dpitch, const uint64_t *dsrc,
                                  size_t spitch, uint64_t r, const uint64_t *ds, int n,
                                  void *stream) {
    size_t nw = words_of(r);
    if (!ddst || !dsrc || !ds || r < 3 || dpitch < nw || spitch < nw || n < 0) {
        gf2t::set_error("gf2t_mulx_xor_slab: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    hipLaunchKernelGGL(k_mulx_xor_slab, dim3(blocks_for(nw, DDF_THREADS), grid_y((unsigned)n)),
                       dim3(DDF_THREADS), 0, (hipStream_t)stream, ddst, (u64)dpitch, dsrc,
                       (u64)spitch, nw, r, ds, (unsigned)n);
    GF2T_CHECK_LAUNCH("k_mulx_xor_slab");
    return GF2T_OK;
}

extern "C" int gf2t_xor_word0_slab(uint64_t *dp, size_t pitch, uint64_t v, int n, void *stream) {
    if (!dp || n < 0) {
        gf2t::set_error("gf2t_xor_word0_slab: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    hipLaunchKernelGGL(k_xor_word0_slab, dim3(blocks_for((size_t)n, 256)), dim3(256), 0,
                       (hipStream_t)stream, dp, (u64)pitch, v, (unsigned)n);
    GF2T_CHECK_LAUNCH("k_xor_word0_slab");
    return GF2T_OK;
}

extern "C" int gf2t_set_one_slab(uint64_t *dp, size_t pitch, size_t nw, int n, void *stream) {
    if (!dp || n < 0 || pitch < nw) {
        gf2t::set_error("gf2t_set_one_slab: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0 || nw == 0) return GF2T_OK;
    hipLaunchKernelGGL(k_set_one_slab, dim3(blocks_for(nw, DDF_THREADS), grid_y((unsigned)n)),
                       dim3(DDF_THREADS), 0, (hipStream_t)stream, dp, (u64)pitch, nw, (unsigned)n);
    GF2T_CHECK_LAUNCH("k_set_one_slab");
    return GF2T_OK;
}

extern "C" int gf2t_copy_slab(uint64_t *ddst, size_t dpitch, const uint64_t *dsrc,
                              size_t spitch, size_t nw, int n, void *stream) {
    if (!ddst || !dsrc || n < 0 || dpitch < nw || spitch < nw) {
        gf2t::set_error("gf2t_copy_slab: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0 || nw == 0) return GF2T_OK;
    hipLaunchKernelGGL(k_copy_slab, dim3(blocks_for(nw, DDF_THREADS), grid_y((unsigned)n)),
                       dim3(DDF_THREADS), 0, (hipStream_t)stream, ddst, (u64)dpitch, dsrc,
                       (u64)spitch, nw, (unsigned)n);
    GF2T_CHECK_LAUNCH("k_copy_slab");
    return GF2T_OK;
}

extern "C" int gf2t_horner_slab(const uint64_t *dh, size_t slot_pitch, int m, uint64_t r,
                                const uint64_t *ds, uint64_t *dc, size_t cpitch, int n,
                                void *stream) {
    size_t nw = words_of(r);
    if (!dh || !dc || !ds || r < 3 || m < 1 || m > DDF_MAX_M || 2 * (u64)m >= r || slot_pitch < nw ||
        cpitch < nw || n < 0) {
        gf2t::set_error("gf2t_horner_slab: bad arguments (1 <= m <= %d, 2m < r)", DDF_MAX_M);
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    hipLaunchKernelGGL(k_horner_slab, dim3(blocks_for(nw, DDF_THREADS), grid_y((unsigned)n)),
                       dim3(DDF_THREADS), 0, (hipStream_t)stream, dh, (u64)slot_pitch,
                       (u64)slot_pitch * (u64)m, m, nw, r, ds, dc, (u64)cpitch, (unsigned)n);
    GF2T_CHECK_LAUNCH("k_horner_slab");
    return GF2T_OK;
}

// h-table of n residues at block start kstart: the Pascal build of
// e_(j+1)(x, x^2, ..., x^(2^(m-1))) (each new variable is the Frobenius
// image of the previous one: e_i(x, Y^2) = e_i(Y)^2 + x e_(i-1)(Y)^2),
// then kstart Frobenius steps (TsGPU::init_h_at, src/tsfactor.cu:
// 345-365).  dh_alt is a second table of the same size used as scratch;
// the result ends in dh.
extern "C" int gf2t_ddf_init_slab(uint64_t *dh, uint64_t *dh_alt, size_t slot_pitch, int m,
                                  uint64_t r, const uint64_t *ds, uint64_t kstart, int n,
                                  void *stream) {
    size_t nw = words_of(r);
    if (!dh || !dh_alt || dh == dh_alt || !ds || r < 3 || m < 1 || m > DDF_MAX_M ||
        slot_pitch < nw || n < 0) {
        gf2t::set_error("gf2t_ddf_init_slab: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    hipStream_t st = (hipStream_t)stream;
    const size_t rp = slot_pitch * (size_t)m;
    const dim3 gw(blocks_for(nw, DDF_THREADS), grid_y((unsigned)n));
    const dim3 gp(blocks_for((nw + 1) / 2, DDF_THREADS), grid_y((unsigned)n));
    int rc = gf2t_memset(dh, 0, rp * (size_t)n * 8, stream);
    if (rc) return rc;
    // slot j's current value lives in dh or in the same slot of dh_alt
    // (alt[j]): a squaring writes the other table, so no copies back
    std::vector<char> alt(m, 0);
    auto slot = [&](int j) { return (alt[j] ? dh_alt : dh) + (size_t)j * slot_pitch; };
    for (int j = 1; j <= m; j++) {
        for (int i = j - 1; i >= 0; i--) {
            if (i > 0) {
                // h[i-1] <- h[i-1]^2, then h[i] ^= x h[i-1]
                u64 *prev = slot(i - 1);
                alt[i - 1] ^= 1;
                u64 *sq = slot(i - 1);
                if (sq_w12_ok(r))
                    hipLaunchKernelGGL(k_square_fold_slab_g<true>, gp, dim3(DDF_THREADS), 0, st,
                                       prev, sq, (u64)rp, r, ds, (unsigned)n, 1u);
                else
                    hipLaunchKernelGGL(k_square_fold_slab_g<false>, gp, dim3(DDF_THREADS), 0, st,
                                       prev, sq, (u64)rp, r, ds, (unsigned)n, 1u);
                hipLaunchKernelGGL(k_mulx_xor_slab, gw, dim3(DDF_THREADS), 0, st, slot(i), (u64)rp,
                                   sq, (u64)rp, nw, r, ds, (unsigned)n);
            } else {
                hipLaunchKernelGGL(k_xor_word0_slab, dim3(blocks_for((size_t)n, 256)), dim3(256),
                                   0, st, slot(0), (u64)rp, 2ULL, (unsigned)n);
            }
        }
    }
    for (int j = 0; j < m; j++)   // the slots whose value ended in dh_alt
        if (alt[j])
            hipLaunchKernelGGL(k_copy_slab, gw, dim3(DDF_THREADS), 0, st,
                               dh + (size_t)j * slot_pitch, (u64)rp,
                               dh_alt + (size_t)j * slot_pitch, (u64)rp, nw, (unsigned)n);
    GF2T_CHECK_LAUNCH("gf2t_ddf_init_slab");
    u64 *cur = dh, *oth = dh_alt;
    const bool two = r >= (1u << 18) && gf2t_host::square_fuse_enabled();
    for (u64 t = 0; t < kstart;) {
        if (two && kstart - t >= 2) {
            rc = gf2t_square2_fold_slab_g(cur, oth, slot_pitch, r, ds, (size_t)n * m, m, stream);
            t += 2;
        } else {
            rc = gf2t_square_fold_slab_g(cur, oth, slot_pitch, r, ds, (size_t)n * m, m, stream);
            t += 1;
        }
        if (rc) return rc;
        std::swap(cur, oth);
    }
    if (cur != dh) return gf2t_memcpy_d2d(dh, cur, rp * (size_t)n * 8, stream);
    return GF2T_OK;
}
