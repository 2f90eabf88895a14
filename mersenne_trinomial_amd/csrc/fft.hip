// fft.hip -- (c) Cantor additive FFT over GF(2^64) and GF(2)[x] products
// on gfx950.
//
// Transform (same evaluation points and natural output order as the
// reference's CantorFFT, src/gf2_cantor_engine.h:125-188): output i is
// the chunk polynomial evaluated at sum_{bits k of i} b[k] of the Cantor
// basis.  In index-bit terms the reference's in-place algorithm is a
// sequence of small linear maps:
//   Taylor level e  (bits e, e+1):  f[..10..] ^= f[..11..]; f[..01..] ^= f[..10..]
//   forward Taylor cascade: for L = 0..m-2: for e = m-2 down to L: T(e)
//   butterfly depth L (bit L): e1 = g + W(t) h, e2 = e1 + h with
//   W(t) = sum_{bits k of t} b[k+1], t = index >> (L+1)
//   forward butterflies: L = m-1 down to 0.
// Each map touches one or two index bits, so a tile that holds every
// element along a set of "local" index bits can apply, in LDS, every map
// whose bits are local.  A host scheduler packs the sequence into as few
// HBM passes as possible, moving a map earlier only past maps it
// commutes with (disjoint bits, and a butterfly's twiddle bits untouched).
// The pass's local bits are the low c bits (for coalesced 128-byte
// segments) plus a window [a, a+w) of higher bits.
//
// Twiddles are linear in t, so each block builds four 256-entry byte
// tables in LDS (W(t) = T0[t&255]^T1[t>>8&255]^...) instead of reading an
// n/2-entry table from HBM.
#include "gf2t_internal.hpp"
#include "../../include/gf2t.h"
#include <vector>
#include <map>
#include <mutex>
#include <condition_variable>
#include <cstring>
#include <algorithm>

namespace {

constexpr int FFT_MAX_OPS = 64;
constexpr int FFT_THREADS = 512;
constexpr int FFT_TILE_BITS = 13;   // 8192 elements = 64 KiB of LDS
constexpr int FFT_LOW_BITS = 4;     // 16 contiguous elements = 128 bytes

enum : uint8_t { OP_TF = 0, OP_TI = 1, OP_BF = 2, OP_BI = 3 };

struct Op {
    uint8_t kind, bit;
};

struct PassDesc {
    int m;        // transform size exponent
    int c, a, w;  // local bits: [0,c) and [a, a+w); tile = 2^(c+w)
    int nops;
    uint8_t kind[FFT_MAX_OPS];
    uint8_t lbit[FFT_MAX_OPS];   // local bit of the op (low bit of a Taylor pair)
    uint8_t gbit[FFT_MAX_OPS];   // global bit (butterfly depth L)
};

__constant__ u64 c_basis[64];

struct TwTables {
    u64 t[4][256];
};

GF2T_DEV u64 twiddle(const TwTables &tw, u64 t) {
    return tw.t[0][t & 255] ^ tw.t[1][(t >> 8) & 255] ^ tw.t[2][(t >> 16) & 255] ^
           tw.t[3][(t >> 24) & 255];
}

GF2T_DEV u64 global_index(const PassDesc &pd, u64 li, u64 rest_lo, u64 rest_hi) {
    u64 cm = (1ULL << pd.c) - 1, wm = (1ULL << pd.w) - 1;
    return (li & cm) | (rest_lo << pd.c) | (((li >> pd.c) & wm) << pd.a) |
           (rest_hi << (pd.a + pd.w));
}

// pack: element k of the transform input = 32-bit chunk k of the bit
// array (zero beyond nchunks; last chunk masked to `lastbits`)
GF2T_DEV u64 packed_chunk(const u32 *__restrict__ src, u64 nchunks, u32 lastmask, u64 k) {
    if (k >= nchunks) return 0;
    u32 v = src[k];
    if (k == nchunks - 1) v &= lastmask;
    return v;
}

// blockIdx.y selects one transform of a batch: element y lives at
// f + y*f_stride and packs from pack_src + y*pack_stride (u32 units);
// pack_xor0 is XORed into element 0 (adds x to the packed operand).
__global__ __launch_bounds__(FFT_THREADS) void k_fft_pass(
        u64 *__restrict__ f, const PassDesc pd, const u32 *__restrict__ pack_src,
        u64 pack_nchunks, u32 pack_lastmask, u64 f_stride, u64 pack_stride, u32 pack_xor0) {
    extern __shared__ u64 lds[];
    f += (u64)blockIdx.y * f_stride;
    if (pack_src) pack_src += (u64)blockIdx.y * pack_stride;
    TwTables &tw = *reinterpret_cast<TwTables *>(lds);
    u64 *tile = lds + 4 * 256;
    const int T = pd.c + pd.w;
    const u64 tsize = 1ULL << T;
    const u64 blk = blockIdx.x;
    const int nlo = pd.a - pd.c;
    const u64 rest_lo = blk & ((1ULL << nlo) - 1);
    const u64 rest_hi = blk >> nlo;

    // twiddle byte tables: T_k[v] = sum_{bits j of v} b[8k + j + 1]
    for (int e = threadIdx.x; e < 1024; e += FFT_THREADS) {
        int k = e >> 8, v = e & 255;
        u64 acc = 0;
        for (int j = 0; j < 8; j++)
            if ((v >> j) & 1) {
                int bi = 8 * k + j + 1;
                acc ^= bi < 64 ? c_basis[bi] : 0;
            }
        tw.t[k][v] = acc;
    }
    // load the tile (low c bits contiguous in memory)
    for (u64 li = threadIdx.x; li < tsize; li += FFT_THREADS) {
        u64 g = global_index(pd, li, rest_lo, rest_hi);
        u64 v = pack_src ? packed_chunk(pack_src, pack_nchunks, pack_lastmask, g) : f[g];
        if (pack_src && g == 0) v ^= pack_xor0;
        tile[li] = v;
    }
    __syncthreads();

    for (int o = 0; o < pd.nops; o++) {
        const int kind = pd.kind[o];
        const int l = pd.lbit[o];
        const u64 lo_mask = (1ULL << l) - 1;
        if (kind == OP_TF || kind == OP_TI) {
            const u64 st = 1ULL << l;
            for (u64 qd = threadIdx.x; qd < (tsize >> 2); qd += FFT_THREADS) {
                u64 base = (qd & lo_mask) | ((qd >> l) << (l + 2));
                u64 x01 = tile[base + st], x10 = tile[base + 2 * st], x11 = tile[base + 3 * st];
                if (kind == OP_TF) { x10 ^= x11; x01 ^= x10; }
                else { x01 ^= x10; x10 ^= x11; }
                tile[base + st] = x01;
                tile[base + 2 * st] = x10;
            }
        } else {
            const int L = pd.gbit[o];
            const u64 st = 1ULL << l;
            for (u64 pr = threadIdx.x; pr < (tsize >> 1); pr += FFT_THREADS) {
                u64 i1 = (pr & lo_mask) | ((pr >> l) << (l + 1));
                u64 g1 = global_index(pd, i1, rest_lo, rest_hi);
                u64 w = twiddle(tw, g1 >> (L + 1));
                u64 e1 = tile[i1], e2 = tile[i1 + st];
                if (kind == OP_BF) {
                    e1 ^= gf64_mul(w, e2);
                    e2 ^= e1;
                } else {
                    e2 ^= e1;
                    e1 ^= gf64_mul(w, e2);
                }
                tile[i1] = e1;
                tile[i1 + st] = e2;
            }
        }
        __syncthreads();
    }
    for (u64 li = threadIdx.x; li < tsize; li += FFT_THREADS)
        f[global_index(pd, li, rest_lo, rest_hi)] = tile[li];
}

__global__ void k_pointwise(u64 *__restrict__ fa, const u64 *__restrict__ fb, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) fa[i] = gf64_mul(fa[i], fb[i]);
}

// overlap-add in gather form: output word w collects chunks 2w-1, 2w, 2w+1
// (blockIdx.y: batch element, F + y*f_stride -> out + y*o_stride)
__global__ void k_overlap_add(const u64 *__restrict__ F, size_t nchunks, u64 *__restrict__ out,
                              size_t owords, u64 f_stride, u64 o_stride) {
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= owords) return;
    F += (u64)blockIdx.y * f_stride;
    out += (u64)blockIdx.y * o_stride;
    size_t k = 2 * w;
    u64 v = 0;
    if (k < nchunks) v ^= F[k];
    if (k + 1 < nchunks) v ^= F[k + 1] << 32;
    if (w && k - 1 < nchunks) v ^= F[k - 1] >> 32;
    out[w] = v;
}

// ------------------------------------------------------------ scheduler
bool op_conflict(const Op &x, const Op &y) {
    bool tx = x.kind <= OP_TI, ty = y.kind <= OP_TI;
    if (tx && ty) return std::abs((int)x.bit - (int)y.bit) <= 1;
    if (!tx && !ty) return true;
    const Op &t = tx ? x : y, &b = tx ? y : x;
    return (int)t.bit + 1 >= (int)b.bit;   // shares bit L or a twiddle bit
}

struct Window {
    int c, a, w;
    bool has(int bit) const { return bit < c || (bit >= a && bit < a + w); }
    int loc(int bit) const { return bit < c ? bit : c + (bit - a); }
};

bool op_fits(const Op &o, const Window &W) {
    if (o.kind <= OP_TI) return W.has(o.bit) && W.has(o.bit + 1);
    return W.has(o.bit);
}

std::vector<Op> op_sequence(int m, bool forward) {
    std::vector<Op> ops;
    if (forward) {
        for (int L = 0; L <= m - 2; L++)
            for (int e = m - 2; e >= L; e--) ops.push_back({OP_TF, (uint8_t)e});
        for (int L = m - 1; L >= 0; L--) ops.push_back({OP_BF, (uint8_t)L});
    } else {
        for (int L = 0; L <= m - 1; L++) ops.push_back({OP_BI, (uint8_t)L});
        for (int L = m - 2; L >= 0; L--)
            for (int e = L; e <= m - 2; e++) ops.push_back({OP_TI, (uint8_t)e});
    }
    return ops;
}

std::vector<PassDesc> schedule(int m, bool forward) {
    std::vector<Op> rem = op_sequence(m, forward);
    std::vector<PassDesc> passes;
    int T = std::min(m, FFT_TILE_BITS);
    int c = std::min(FFT_LOW_BITS, T);
    int w = T - c;
    while (!rem.empty()) {
        std::vector<Window> cands;
        if (m <= T) cands.push_back({m, m, 0});
        else
            for (int a = c; a + w <= m; a++) cands.push_back({c, a, w});
        size_t best_n = 0;
        Window bestW = cands[0];
        std::vector<Op> best_take, best_skip;
        for (const Window &W : cands) {
            std::vector<Op> take, skip;
            for (const Op &o : rem) {
                bool ok = op_fits(o, W) && (int)take.size() < FFT_MAX_OPS;
                if (ok)
                    for (const Op &s : skip)
                        if (op_conflict(o, s)) { ok = false; break; }
                (ok ? take : skip).push_back(o);
            }
            if (take.size() > best_n) {
                best_n = take.size();
                bestW = W;
                best_take.swap(take);
                best_skip.swap(skip);
            }
        }
        if (best_n == 0) break;   // cannot happen: some window fits the head op
        PassDesc pd{};
        pd.m = m;
        pd.c = bestW.c;
        pd.a = bestW.a;
        pd.w = bestW.w;
        pd.nops = (int)best_take.size();
        for (int i = 0; i < pd.nops; i++) {
            pd.kind[i] = best_take[i].kind;
            pd.lbit[i] = (uint8_t)bestW.loc(best_take[i].bit);
            pd.gbit[i] = best_take[i].bit;
        }
        passes.push_back(pd);
        rem.swap(best_skip);
    }
    return passes;
}

// --------------------------------------------------------- basis (host)
u64 gf64_mul_host(u64 a, u64 b) { return gf64_mul(a, b); }

// y^2 + y = c: Gauss-Jordan on the 64x64 matrix of T(y) = y^2 + y; of the
// two roots (they differ by 1) return the one with bit 0 clear, which is
// the reference's choice (src/gf2_cantor_engine.h:78-110).
bool solve_T(u64 c, u64 &y) {
    u64 row[64];
    int rhs[64];
    for (int j = 0; j < 64; j++) { row[j] = 0; rhs[j] = (int)((c >> j) & 1); }
    for (int i = 0; i < 64; i++) {
        u64 ei = 1ULL << i, t = gf64_mul_host(ei, ei) ^ ei;
        for (int j = 0; j < 64; j++)
            if ((t >> j) & 1) row[j] |= ei;
    }
    int piv[64];
    for (int i = 0; i < 64; i++) piv[i] = -1;
    int rank = 0;
    for (int col = 63; col >= 0; col--) {
        int p = -1;
        for (int j = rank; j < 64; j++)
            if ((row[j] >> col) & 1) { p = j; break; }
        if (p < 0) continue;
        std::swap(row[p], row[rank]);
        std::swap(rhs[p], rhs[rank]);
        for (int j = 0; j < 64; j++)
            if (j != rank && ((row[j] >> col) & 1)) { row[j] ^= row[rank]; rhs[j] ^= rhs[rank]; }
        piv[col] = rank++;
    }
    for (int j = rank; j < 64; j++)
        if (rhs[j]) return false;
    u64 sol = 0;
    for (int col = 0; col < 64; col++)
        if (piv[col] >= 0 && rhs[piv[col]]) sol |= 1ULL << col;
    if (sol & 1) sol ^= 1;
    if ((gf64_mul_host(sol, sol) ^ sol) != c) return false;
    y = sol;
    return true;
}

struct BasisState {
    std::mutex mu;
    bool built = false;
    u64 b[64];
    int len = 0;
    int uploaded_dev = -1;
} g_basis;

void build_basis_host() {
    std::lock_guard<std::mutex> lk(g_basis.mu);
    if (g_basis.built) return;
    g_basis.b[0] = 1;
    g_basis.len = 1;
    while (g_basis.len < 64) {
        u64 y;
        if (!solve_T(g_basis.b[g_basis.len - 1], y)) break;
        g_basis.b[g_basis.len++] = y;
    }
    g_basis.built = true;
}

int ensure_basis() {
    build_basis_host();
    int dev = 0;
    GF2T_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_basis.mu);
    if (g_basis.uploaded_dev != dev) {
        GF2T_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_basis), g_basis.b, sizeof(g_basis.b)));
        g_basis.uploaded_dev = dev;
    }
    return GF2T_OK;
}

}  // namespace

struct gf2t_fft_plan {
    int m;
    std::vector<PassDesc> fwd, inv;
};

namespace {
std::mutex g_plan_mu;
std::map<int, gf2t_fft_plan *> g_plans;

const gf2t_fft_plan *cached_plan(int m) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(m);
    if (it != g_plans.end()) return it->second;
    gf2t_fft_plan *p = new gf2t_fft_plan{m, schedule(m, true), schedule(m, false)};
    g_plans[m] = p;
    return p;
}

int run_passes(const std::vector<PassDesc> &ps, u64 *df, hipStream_t st, const u32 *pack_src,
               u64 nchunks, u32 lastmask, unsigned nbatch = 1, u64 f_stride = 0,
               u64 pack_stride = 0, u32 pack_xor0 = 0) {
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    std::call_once(lds_once, [] {
        lds_err = hipFuncSetAttribute((const void *)k_fft_pass,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)((4 * 256 + ((size_t)1 << FFT_TILE_BITS)) * sizeof(u64)));
    });
    GF2T_HIP(lds_err);
    for (size_t i = 0; i < ps.size(); i++) {
        const PassDesc &pd = ps[i];
        int T = pd.c + pd.w;
        unsigned blocks = (unsigned)(1ULL << (pd.m - T));
        size_t lds = (4 * 256 + ((size_t)1 << T)) * sizeof(u64);
        const u32 *src = (i == 0) ? pack_src : nullptr;
        hipLaunchKernelGGL(k_fft_pass, dim3(blocks, nbatch), dim3(FFT_THREADS), lds, st, df, pd,
                           src, nchunks, lastmask, f_stride, pack_stride, pack_xor0);
    }
    GF2T_CHECK_LAUNCH("k_fft_pass");
    return GF2T_OK;
}

int fft_m_for(u64 nchunks_total) {
    int m = 1;
    while (((u64)1 << m) < nchunks_total) m++;
    return m;
}

constexpr int FFT_MAX_M = 27;
}  // namespace

extern "C" int gf2t_cantor_basis(uint64_t *hbasis, int *len) {
    build_basis_host();
    if (hbasis) std::memcpy(hbasis, g_basis.b, sizeof(g_basis.b));
    if (len) *len = g_basis.len;
    return GF2T_OK;
}

extern "C" int gf2t_fft_plan_create(int m, gf2t_fft_plan **plan) {
    if (!plan || m < 1 || m > FFT_MAX_M) {
        gf2t::set_error("gf2t_fft_plan_create: need 1 <= m <= %d", FFT_MAX_M);
        return GF2T_EINVAL;
    }
    build_basis_host();
    if (g_basis.len < m) {
        gf2t::set_error("gf2t_fft_plan_create: basis length %d < m", g_basis.len);
        return GF2T_EINTERNAL;
    }
    *plan = new gf2t_fft_plan{m, schedule(m, true), schedule(m, false)};
    return GF2T_OK;
}

extern "C" int gf2t_fft_plan_destroy(gf2t_fft_plan *plan) {
    delete plan;
    return GF2T_OK;
}

extern "C" int gf2t_fft_plan_passes(const gf2t_fft_plan *plan, int *fwd_passes, int *inv_passes) {
    if (!plan) return GF2T_EINVAL;
    if (fwd_passes) *fwd_passes = (int)plan->fwd.size();
    if (inv_passes) *inv_passes = (int)plan->inv.size();
    return GF2T_OK;
}

extern "C" int gf2t_fft_fwd(const gf2t_fft_plan *plan, uint64_t *df, void *stream) {
    if (!plan || !df) return GF2T_EINVAL;
    int rc = ensure_basis();
    if (rc) return rc;
    return run_passes(plan->fwd, df, (hipStream_t)stream, nullptr, 0, 0);
}

extern "C" int gf2t_fft_inv(const gf2t_fft_plan *plan, uint64_t *df, void *stream) {
    if (!plan || !df) return GF2T_EINVAL;
    int rc = ensure_basis();
    if (rc) return rc;
    return run_passes(plan->inv, df, (hipStream_t)stream, nullptr, 0, 0);
}

extern "C" size_t gf2t_mul_out_words(uint64_t abits, uint64_t bbits) {
    return words_of(abits + bbits) + 1;
}

extern "C" size_t gf2t_mul_workspace_bytes(uint64_t abits, uint64_t bbits) {
    u64 ca = (abits + 31) >> 5, cb = (bbits + 31) >> 5;
    if (!ca || !cb) return 0;
    int m = fft_m_for(ca + cb - 1);
    size_t n = (size_t)1 << m;
    return (2 * n + gf2t_mul_out_words(abits, bbits) + 8) * sizeof(u64);
}

namespace {
int mul_core(const u64 *da, u64 abits, const u64 *db, u64 bbits, u64 *dprod, void *dwork,
             hipStream_t st) {
    u64 ca = (abits + 31) >> 5, cb = (bbits + 31) >> 5;
    size_t ow = gf2t_mul_out_words(abits, bbits);
    if (!ca || !cb) {
        GF2T_HIP(hipMemsetAsync(dprod, 0, ow * 8, st));
        return GF2T_OK;
    }
    int m = fft_m_for(ca + cb - 1);
    if (m > FFT_MAX_M) {
        gf2t::set_error("gf2t_mul: product needs 2^%d points (max 2^%d)", m, FFT_MAX_M);
        return GF2T_ERANGE;
    }
    int rc = ensure_basis();
    if (rc) return rc;
    const gf2t_fft_plan *plan = cached_plan(m);
    size_t n = (size_t)1 << m;
    u64 *FA = (u64 *)dwork, *FB = FA + n;
    u32 ma = (abits & 31) ? (u32)((1ULL << (abits & 31)) - 1) : 0xFFFFFFFFu;
    u32 mb = (bbits & 31) ? (u32)((1ULL << (bbits & 31)) - 1) : 0xFFFFFFFFu;
    rc = run_passes(plan->fwd, FA, st, (const u32 *)da, ca, ma);
    if (rc) return rc;
    rc = run_passes(plan->fwd, FB, st, (const u32 *)db, cb, mb);
    if (rc) return rc;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_pointwise, dim3(blocks), dim3(256), 0, st, FA, FB, n);
    rc = run_passes(plan->inv, FA, st, nullptr, 0, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_overlap_add, dim3((unsigned)((ow + 255) / 256)), dim3(256), 0, st, FA,
                       (size_t)(ca + cb - 1), dprod, ow, (u64)0, (u64)0);
    GF2T_CHECK_LAUNCH("k_overlap_add");
    return GF2T_OK;
}
}  // namespace

extern "C" int gf2t_mul(const uint64_t *da, uint64_t abits, const uint64_t *db, uint64_t bbits,
                        uint64_t *dout, void *dwork, void *stream) {
    if (!da || !db || !dout || (!dwork && abits && bbits)) {
        gf2t::set_error("gf2t_mul: null pointer");
        return GF2T_EINVAL;
    }
    return mul_core(da, abits, db, bbits, dout, dwork, (hipStream_t)stream);
}

extern "C" int gf2t_fold(const uint64_t *dsrc, uint64_t src_bits, uint64_t *dout, uint64_t r,
                         uint64_t s, void *stream);

extern "C" int gf2t_mulmod(const uint64_t *da, uint64_t abits, const uint64_t *db,
                           uint64_t bbits, uint64_t r, uint64_t s, uint64_t *dout, void *dwork,
                           void *stream) {
    if (!da || !db || !dout || !dwork) {
        gf2t::set_error("gf2t_mulmod: null pointer");
        return GF2T_EINVAL;
    }
    if (abits + bbits > 2 * r) {
        gf2t::set_error("gf2t_mulmod: product of %llu+%llu bits exceeds 2r-1",
                        (unsigned long long)abits, (unsigned long long)bbits);
        return GF2T_EINVAL;
    }
    u64 ca = (abits + 31) >> 5, cb = (bbits + 31) >> 5;
    size_t n = (ca && cb) ? ((size_t)1 << fft_m_for(ca + cb - 1)) : 0;
    u64 *prod = (u64 *)dwork + 2 * n;
    int rc = mul_core(da, abits, db, bbits, prod, dwork, (hipStream_t)stream);
    if (rc) return rc;
    u64 pbits = abits + bbits ? abits + bbits - 1 : 0;
    if (pbits < 1) pbits = 1;
    return gf2t_fold(prod, pbits, dout, r, s, stream);
}

// ------------------------------------------------------------------ slabs
namespace {
__global__ void k_fold_slab(const u64 *__restrict__ p, size_t pw, u64 p_stride, u64 L,
                            u64 *__restrict__ out, u64 o_stride, size_t nw, u64 r,
                            const u64 *__restrict__ svec) {
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const u64 s = svec[blockIdx.y];
    p += (u64)blockIdx.y * p_stride;
    out += (u64)blockIdx.y * o_stride;
    out[w] = fold_word(p, pw, L, r, s, w);
}
}  // namespace

extern "C" size_t gf2t_mulmod_slab_workspace_bytes(uint64_t r, int n) {
    u64 ca = (r + 31) >> 5;
    int m = fft_m_for(2 * ca - 1);
    size_t per = 2 * ((size_t)1 << m) + 2 * words_of(r) + 2;
    return per * (size_t)(n > 0 ? n : 0) * sizeof(u64);
}

extern "C" int gf2t_mulmod_slab(const uint64_t *da, size_t apitch, const uint64_t *db,
                                size_t bpitch, int b_add_x, uint64_t r, const uint64_t *ds,
                                uint64_t *dout, size_t opitch, int n, void *dwork, void *stream) {
    if (!da || !db || !dout || !ds || !dwork || n < 0 || r < 3) {
        gf2t::set_error("gf2t_mulmod_slab: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    size_t nw = words_of(r);
    if (apitch < nw || bpitch < nw || opitch < nw || n > 65535) {
        gf2t::set_error("gf2t_mulmod_slab: pitch below nw or more than 65535 residues");
        return GF2T_EINVAL;
    }
    hipStream_t st = (hipStream_t)stream;
    int rc = ensure_basis();
    if (rc) return rc;
    u64 ca = (r + 31) >> 5;
    int m = fft_m_for(2 * ca - 1);
    if (m > FFT_MAX_M) { gf2t::set_error("gf2t_mulmod_slab: r too large"); return GF2T_ERANGE; }
    const gf2t_fft_plan *plan = cached_plan(m);
    size_t N = (size_t)1 << m;
    u64 *FA = (u64 *)dwork, *FB = FA + N * n, *prod = FB + N * n;
    size_t ppitch = 2 * nw + 2;
    u32 lm = (r & 31) ? (u32)((1ULL << (r & 31)) - 1) : 0xFFFFFFFFu;
    rc = run_passes(plan->fwd, FA, st, (const u32 *)da, ca, lm, (unsigned)n, N, 2 * apitch, 0);
    if (rc) return rc;
    rc = run_passes(plan->fwd, FB, st, (const u32 *)db, ca, lm, (unsigned)n, N, 2 * bpitch,
                    b_add_x ? 2u : 0u);
    if (rc) return rc;
    size_t tot = N * (size_t)n;
    unsigned blocks = (unsigned)std::min<size_t>((tot + 255) / 256, 65536);
    hipLaunchKernelGGL(k_pointwise, dim3(blocks), dim3(256), 0, st, FA, FB, tot);
    rc = run_passes(plan->inv, FA, st, nullptr, 0, 0, (unsigned)n, N, 0, 0);
    if (rc) return rc;
    size_t ow = 2 * nw;   // product of two r-bit operands: 2r-1 bits
    hipLaunchKernelGGL(k_overlap_add, dim3((unsigned)((ow + 255) / 256), n), dim3(256), 0, st, FA,
                       (size_t)(2 * ca - 1), prod, ow, (u64)N, (u64)ppitch);
    hipLaunchKernelGGL(k_fold_slab, dim3((unsigned)((nw + 255) / 256), n), dim3(256), 0, st, prod,
                       ow, (u64)ppitch, 2 * r - 1, dout, (u64)opitch, nw, r, ds);
    GF2T_CHECK_LAUNCH("gf2t_mulmod_slab");
    return GF2T_OK;
}

// --------------------------------------------- host-buffer product service
// The device side of the reference's GpuMulHook (TsGpuMul, src/tsfactor.cu:
// 467-607; interface src/gf2_gcd.h:155-172): host vectors in and out, one
// mutex-guarded context per process with a private stream and device
// buffers grown on demand (never shrunk).
namespace {
__global__ void k_mat2_pointwise(u64 *__restrict__ f0, const u64 *__restrict__ f1,
                                 u64 *__restrict__ f2, const u64 *__restrict__ f3,
                                 const u64 *__restrict__ fa, const u64 *__restrict__ fb, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        u64 a = fa[i], b = fb[i];
        f0[i] = gf64_mul(f0[i], a) ^ gf64_mul(f1[i], b);
        f2[i] = gf64_mul(f2[i], a) ^ gf64_mul(f3[i], b);
    }
}

// A pool of contexts, each with its own non-blocking stream and device
// buffer, so products requested by different host threads (the gcds of
// different s values) run concurrently on the device.
constexpr int HM_CONTEXTS = 8;
struct HostMulCtx {
    hipStream_t st = nullptr;
    u64 *dbuf = nullptr;
    size_t dcap = 0;   // words
    bool busy = false;
};
struct HostMulPool {
    std::mutex mu;
    std::condition_variable cv;
    HostMulCtx ctx[HM_CONTEXTS];
} g_hm;

// RAII: a free context for the duration of one product
struct HmLease {
    HostMulCtx *c = nullptr;
    HmLease() {
        std::unique_lock<std::mutex> lk(g_hm.mu);
        for (;;) {
            for (auto &x : g_hm.ctx)
                if (!x.busy) { x.busy = true; c = &x; return; }
            g_hm.cv.wait(lk);
        }
    }
    ~HmLease() {
        {
            std::lock_guard<std::mutex> lk(g_hm.mu);
            c->busy = false;
        }
        g_hm.cv.notify_one();
    }
};

int hm_reserve(HostMulCtx &c, size_t words) {
    if (!c.st) GF2T_HIP(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    if (words <= c.dcap) return GF2T_OK;
    if (c.dbuf) GF2T_HIP(hipFree(c.dbuf));
    c.dbuf = nullptr;
    c.dcap = 0;
    size_t want = words + words / 4;
    if (hipMalloc(&c.dbuf, want * 8) != hipSuccess) {
        (void)hipGetLastError();
        gf2t::set_error("gf2t host product service: cannot allocate %zu bytes", want * 8);
        return GF2T_ENOMEM;
    }
    c.dcap = want;
    return GF2T_OK;
}

u32 last_mask(u64 bits) { return (bits & 31) ? (u32)((1ULL << (bits & 31)) - 1) : 0xFFFFFFFFu; }
}  // namespace

extern "C" int gf2t_mul_host(const uint64_t *ha, uint64_t abits, const uint64_t *hb,
                             uint64_t bbits, uint64_t *hout) {
    if ((!ha && abits) || (!hb && bbits) || !hout) {
        gf2t::set_error("gf2t_mul_host: null pointer");
        return GF2T_EINVAL;
    }
    size_t ow = gf2t_mul_out_words(abits, bbits);
    if (!abits || !bbits) {
        std::memset(hout, 0, ow * 8);
        return GF2T_OK;
    }
    HmLease lease;
    HostMulCtx &hc = *lease.c;
    size_t aw = words_of(abits), bw = words_of(bbits);
    size_t ww = (gf2t_mul_workspace_bytes(abits, bbits) + 7) / 8;
    int rc = hm_reserve(hc, aw + bw + ow + ww + 64);   // + room to align the workspace
    if (rc) return rc;
    u64 *da = hc.dbuf, *db = da + aw, *dout = db + bw, *dw = dout + ow;
    dw = (u64 *)(((uintptr_t)dw + 255) & ~(uintptr_t)255);
    hipStream_t st = hc.st;
    GF2T_HIP(hipMemcpyAsync(da, ha, aw * 8, hipMemcpyHostToDevice, st));
    GF2T_HIP(hipMemcpyAsync(db, hb, bw * 8, hipMemcpyHostToDevice, st));
    rc = mul_core(da, abits, db, bbits, dout, dw, st);
    if (rc) return rc;
    GF2T_HIP(hipMemcpyAsync(hout, dout, ow * 8, hipMemcpyDeviceToHost, st));
    GF2T_HIP(hipStreamSynchronize(st));
    return GF2T_OK;
}

extern "C" int gf2t_mat2_host(const uint64_t *const hm[4], const uint64_t mbits[4],
                              const uint64_t *ha, uint64_t abits, const uint64_t *hb,
                              uint64_t bbits, uint64_t *hoa, uint64_t *hob, size_t outw) {
    if (!hm || !mbits || !hoa || !hob) {
        gf2t::set_error("gf2t_mat2_host: null pointer");
        return GF2T_EINVAL;
    }
    u64 mb = 0;
    for (int i = 0; i < 4; i++) {
        if (!hm[i] && mbits[i]) { gf2t::set_error("gf2t_mat2_host: null entry"); return GF2T_EINVAL; }
        mb = std::max<u64>(mb, mbits[i]);
    }
    u64 vb = std::max(abits, bbits);
    if ((!ha && abits) || (!hb && bbits)) { gf2t::set_error("gf2t_mat2_host: null operand"); return GF2T_EINVAL; }
    if (outw < gf2t_mul_out_words(mb, vb)) {
        gf2t::set_error("gf2t_mat2_host: outw %zu below %zu", outw, gf2t_mul_out_words(mb, vb));
        return GF2T_EINVAL;
    }
    std::memset(hoa, 0, outw * 8);
    std::memset(hob, 0, outw * 8);
    if (!mb || !vb) return GF2T_OK;
    u64 cm = (mb + 31) >> 5, cv = (vb + 31) >> 5;
    int m = fft_m_for(cm + cv - 1);
    if (m > FFT_MAX_M) { gf2t::set_error("gf2t_mat2_host: product too large"); return GF2T_ERANGE; }
    int rc = ensure_basis();
    if (rc) return rc;
    HmLease lease;
    HostMulCtx &hc = *lease.c;
    const gf2t_fft_plan *plan = cached_plan(m);
    size_t n = (size_t)1 << m, ow = gf2t_mul_out_words(mb, vb);
    size_t in_w[6] = {words_of(mbits[0]), words_of(mbits[1]), words_of(mbits[2]),
                      words_of(mbits[3]), words_of(abits), words_of(bbits)};
    size_t stage = 0;
    for (size_t w : in_w) stage += w + 4;
    rc = hm_reserve(hc, 6 * n + 2 * ow + stage + 64);
    if (rc) return rc;
    hipStream_t st = hc.st;
    u64 *F[6];
    for (int i = 0; i < 6; i++) F[i] = hc.dbuf + (size_t)i * n;
    u64 *dout = hc.dbuf + 6 * n, *dst = dout + 2 * ow;
    const u64 *src[6] = {hm[0], hm[1], hm[2], hm[3], ha, hb};
    const u64 bits[6] = {mbits[0], mbits[1], mbits[2], mbits[3], abits, bbits};
    for (int i = 0; i < 6; i++) {
        u64 c = (bits[i] + 31) >> 5;
        if (!c) {
            GF2T_HIP(hipMemsetAsync(F[i], 0, n * 8, st));
            continue;
        }
        GF2T_HIP(hipMemcpyAsync(dst, src[i], in_w[i] * 8, hipMemcpyHostToDevice, st));
        rc = run_passes(plan->fwd, F[i], st, (const u32 *)dst, c, last_mask(bits[i]));
        if (rc) return rc;
        dst += in_w[i] + 4;
    }
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_mat2_pointwise, dim3(blocks), dim3(256), 0, st, F[0], F[1], F[2], F[3],
                       F[4], F[5], n);
    rc = run_passes(plan->inv, F[0], st, nullptr, 0, 0);
    if (rc) return rc;
    rc = run_passes(plan->inv, F[2], st, nullptr, 0, 0);
    if (rc) return rc;
    size_t nch = (size_t)(cm + cv - 1);
    hipLaunchKernelGGL(k_overlap_add, dim3((unsigned)((ow + 255) / 256)), dim3(256), 0, st, F[0],
                       nch, dout, ow, (u64)0, (u64)0);
    hipLaunchKernelGGL(k_overlap_add, dim3((unsigned)((ow + 255) / 256)), dim3(256), 0, st, F[2],
                       nch, dout + ow, ow, (u64)0, (u64)0);
    GF2T_CHECK_LAUNCH("gf2t_mat2_host");
    GF2T_HIP(hipMemcpyAsync(hoa, dout, ow * 8, hipMemcpyDeviceToHost, st));
    GF2T_HIP(hipMemcpyAsync(hob, dout + ow, ow * 8, hipMemcpyDeviceToHost, st));
    GF2T_HIP(hipStreamSynchronize(st));
    return GF2T_OK;
}
