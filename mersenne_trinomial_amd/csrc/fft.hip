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
#include <atomic>

namespace {

constexpr int FFT_MAX_OPS = 64;
constexpr int FFT_THREADS = 512;    // most threads per block (a tile of 2^T elements uses 2^(T-1))
constexpr int FFT_TILE_BITS = 13;   // largest tile: 8192 elements = 64 KiB of LDS
constexpr int FFT_MIN_TILE_BITS = 6;
constexpr int FFT_LOW_BITS = 4;     // default: 16 contiguous elements = 128 bytes
constexpr int PACK_LIST = 12;   // four matrix entries + four vector pairs

enum : uint8_t { OP_TF = 0, OP_TI = 1, OP_BF = 2, OP_BI = 3 };

struct Op {
    uint8_t kind, bit;
};

struct PassDesc {
    int m;        // transform size exponent
    int c, a, w;  // local bits: [0,c) and [a, a+w); tile = 2^(c+w)
    int nops;
    int nbf;      // butterfly ops among them (0: no twiddle tables needed)
    uint8_t kind[FFT_MAX_OPS];
    uint8_t lbit[FFT_MAX_OPS];   // local bit of the op (low bit of a Taylor pair)
    uint8_t gbit[FFT_MAX_OPS];   // global bit (butterfly depth L)
};

__constant__ u64 c_basis[64];

struct TwTables {
    u64 t[4][256];
};

// Where the first pass of a batch of transforms packs its input from:
// 32-bit chunk k of an operand is element k (zero from nchunks on, the
// last chunk masked to lastmask).  Stride form: batch element y packs
// from src + y*stride.  List form (nlist > 0): element y < nlist packs
// from lsrc[y] with its own chunk count and mask, so operands of
// different lengths share one launch.  xor0 is XORed into element 0
// (adds x to the operand).
struct PackArg {
    const u32 *src;
    u64 stride;
    u64 nchunks;
    u32 lastmask, xor0;
    int nlist;
    const u32 *lsrc[PACK_LIST];
    u64 lnch[PACK_LIST];
    u32 lmask[PACK_LIST];
};

// W(t) for t < 2^32: four independent table reads, no branches (the
// tables past the ones a transform needs hold only their entry 0 = 0,
// which is all that a t below 2^(8 ntab) reads there)
GF2T_DEV u64 twiddle(const TwTables &tw, u64 t) {
    const u32 v = (u32)t;
    return xor3(tw.t[0][v & 255], tw.t[1][(v >> 8) & 255], tw.t[2][(v >> 16) & 255]) ^
           tw.t[3][v >> 24];
}

// the global index of tile element li (li < 2^(c+w)): its low c bits stay,
// its bits [c, c+w) go to [a, a+w), and gbase holds the tile's own bits
GF2T_DEV u64 global_index(const PassDesc &pd, u32 li, u64 gbase) {
    const u32 cm = (1u << pd.c) - 1;
    return gbase | (u64)(li & cm) | ((u64)(li >> pd.c) << pd.a);
}

// pack: element k of the transform input = 32-bit chunk k of the bit
// array (zero beyond nchunks; last chunk masked to `lastbits`)
GF2T_DEV u64 packed_chunk(const u32 *__restrict__ src, u64 nchunks, u32 lastmask, u64 k) {
    if (k >= nchunks) return 0;
    u32 v = src[k];
    if (k == nchunks - 1) v &= lastmask;
    return v;
}

// One pass over the tile `blk` of the transform at f (2^m elements):
// load the tile (packing the input chunks on the first pass), apply the
// pass's ops in LDS, store.  The block's threads stride over the tile;
// ntab byte tables of twiddles cover the twiddle indices t < 2^(8 ntab)
// the transform uses.
GF2T_DEV void fft_pass_tile(u64 *__restrict__ f, const PassDesc &pd, bool pack,
                            const u32 *__restrict__ psrc, u64 pn, u32 pmask, u32 xor0, int ntab,
                            u64 blk, u64 *lds) {
    TwTables &tw = *reinterpret_cast<TwTables *>(lds);
    u64 *tile = lds + 4 * 256;
    const int T = pd.c + pd.w;
    const u32 tsize = 1u << T;
    const unsigned nthr = blockDim.x;
    const int nlo = pd.a - pd.c;
    // the tile's fixed index bits: blk's low nlo bits at [c, a), the rest
    // from a + w up
    const u64 gbase = ((blk & ((1ULL << nlo) - 1)) << pd.c) | ((blk >> nlo) << (pd.a + pd.w));

    // twiddle byte tables: T_k[v] = sum_{bits j of v} b[8k + j + 1]
    if (pd.nbf == 0) ntab = 0;
    for (int e = threadIdx.x; e < 256 * ntab; e += nthr) {
        int k = e >> 8, v = e & 255;
        u64 acc = 0;
        for (int j = 0; j < 8; j++)
            if ((v >> j) & 1) {
                int bi = 8 * k + j + 1;
                acc ^= bi < 64 ? c_basis[bi] : 0;
            }
        tw.t[k][v] = acc;
    }
    if (pd.nbf && threadIdx.x < 4 - ntab) tw.t[3 - threadIdx.x][0] = 0;
    // load the tile (low c bits contiguous in memory)
    for (u32 li = threadIdx.x; li < tsize; li += nthr) {
        const u64 g = global_index(pd, li, gbase);
        u64 v;
        if (pack) {
            v = packed_chunk(psrc, pn, pmask, g);
            if (g == 0) v ^= xor0;
        } else {
            v = f[g];
        }
        tile[li] = v;
    }
    __syncthreads();

    for (int o = 0; o < pd.nops; o++) {
        const int kind = pd.kind[o];
        const int l = pd.lbit[o];
        const u32 lo_mask = (1u << l) - 1;
        if (kind == OP_TF || kind == OP_TI) {
            const u32 st = 1u << l;
            for (u32 qd = threadIdx.x; qd < (tsize >> 2); qd += nthr) {
                const u32 base = (qd & lo_mask) | ((qd >> l) << (l + 2));
                u64 x01 = tile[base + st], x10 = tile[base + 2 * st], x11 = tile[base + 3 * st];
                if (kind == OP_TF) { x10 ^= x11; x01 ^= x10; }
                else { x01 ^= x10; x10 ^= x11; }
                tile[base + st] = x01;
                tile[base + 2 * st] = x10;
            }
        } else {
            const int L = pd.gbit[o];
            const u32 st = 1u << l;
            for (u32 pr = threadIdx.x; pr < (tsize >> 1); pr += nthr) {
                const u32 i1 = (pr & lo_mask) | ((pr >> l) << (l + 1));
                const u64 g1 = global_index(pd, i1, gbase);
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
    for (u32 li = threadIdx.x; li < tsize; li += nthr)
        f[global_index(pd, li, gbase)] = tile[li];
}

// blockIdx.y selects one transform of a batch: element y lives at
// f + y*f_stride.  pack (first pass only) describes the input to pack.
// The block has 2^(T-1) threads (at most FFT_THREADS) for its 2^T-element
// tile.
__global__ __launch_bounds__(FFT_THREADS) void k_fft_pass(u64 *__restrict__ f, const PassDesc pd,
                                                          const PackArg pk, int pack, int ntab,
                                                          u64 f_stride) {
    extern __shared__ u64 lds[];
    const unsigned y = blockIdx.y;
    const u32 *psrc = nullptr;
    u64 pn = 0;
    u32 pmask = 0;
    if (pack) {
        if (pk.nlist) {
            psrc = pk.lsrc[y];
            pn = pk.lnch[y];
            pmask = pk.lmask[y];
        } else {
            psrc = pk.src + (u64)y * pk.stride;
            pn = pk.nchunks;
            pmask = pk.lastmask;
        }
    }
    fft_pass_tile(f + (u64)y * f_stride, pd, pack != 0, psrc, pn, pmask, pk.xor0, ntab,
                  blockIdx.x, lds);
}

// A grouped launch: one pass each of several transforms of different
// sizes (the i-th pass of every product in a batch that has one).  Entry
// e owns blocks [block0, block0 + nblocks * ntrans) of the grid: ntrans
// transforms n words apart, nblocks tiles each; a block finds its entry
// by its index.
struct GroupPass {
    PassDesc pd;
    u64 *f;
    u64 f_stride;
    const u32 *src[2];
    u64 nch[2];
    u32 mask[2];
    int pack, ntab;
    u32 block0, nblocks, ntrans, pad;
};

__global__ __launch_bounds__(FFT_THREADS) void k_fft_pass_grouped(
        const GroupPass *__restrict__ gp, int nent) {
    extern __shared__ u64 lds[];
    const u32 b = blockIdx.x;
    int e = 0;
    while (e + 1 < nent && gp[e + 1].block0 <= b) e++;
    const GroupPass &G = gp[e];
    const u32 local = b - G.block0;
    const u32 y = local / G.nblocks, x = local - y * G.nblocks;
    fft_pass_tile(G.f + (u64)y * G.f_stride, G.pd, G.pack != 0, G.pack ? G.src[y] : nullptr,
                  G.pack ? G.nch[y] : 0, G.pack ? G.mask[y] : 0, 0, G.ntab, x, lds);
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

// Pack the op sequence of a 2^m-point transform into passes over tiles of
// 2^min(m, tile_bits) elements.
std::vector<PassDesc> schedule(int m, bool forward, int tile_bits, int low_bits) {
    std::vector<Op> rem = op_sequence(m, forward);
    std::vector<PassDesc> passes;
    int T = std::min(m, tile_bits);
    // at least two window bits, so that every Taylor step fits some window
    int c = std::min(low_bits, std::max(T - 2, 1));
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
        pd.nbf = 0;
        for (int i = 0; i < pd.nops; i++) {
            pd.nbf += best_take[i].kind >= OP_BF;
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
    int m, T;   // 2^m points, tiles of 2^T elements
    std::vector<PassDesc> fwd, inv;
};

namespace {
constexpr int FFT_MAX_M = 27;
std::mutex g_plan_mu;
std::map<int, gf2t_fft_plan *> g_plans;   // key (m * 32 + T) * 8 + low bits

// Tile bits per transform size: whole 8192-element tiles from 2^19 points
// up; below that smaller tiles, so that a transform still spreads over
// tens of workgroups (a 2^15-point transform in 8192-element tiles
// occupies four CUs, each running 15 dependent butterfly levels).
// Overridden per size by gf2t_tune("fft_tile_<m>", T), which is how the
// autotuner installs what it measured.
int default_tile_bits(int m) {
    if (m >= 19) return FFT_TILE_BITS;
    if (m >= 14) return 11;
    return std::min(m, FFT_TILE_BITS);
}
std::atomic<int> g_tile_bits[FFT_MAX_M + 1];   // 0: default_tile_bits

int tile_bits_for(int m) {
    int T = g_tile_bits[m].load(std::memory_order_relaxed);
    return T ? T : default_tile_bits(m);
}

// Threads of a pass over a 2^T-element tile: one per `ept` elements
// (gf2t_tune "fft_ept": 2, 4, 8 or 16; default 2), between 64 and
// FFT_THREADS.  Fewer threads per tile leave room for more blocks of
// other launches on a CU (concurrent products of small transforms).
std::atomic<int> g_fft_ept{2};

// Low index bits of every pass's tile (gf2t_tune "fft_low_bits", 4..6):
// runs of 2^c contiguous elements per HBM access of a tile
std::atomic<int> g_fft_low{FFT_LOW_BITS};
unsigned threads_for_tile(int T) {
    const u64 ept = (u64)g_fft_ept.load(std::memory_order_relaxed);
    return (unsigned)std::max<u64>(64, std::min<u64>(FFT_THREADS, (1ULL << (T > 0 ? T : 0)) / ept));
}

const gf2t_fft_plan *cached_plan_T(int m, int T) {
    T = std::max(FFT_MIN_TILE_BITS, std::min(T, FFT_TILE_BITS));
    const int c = g_fft_low.load(std::memory_order_relaxed);
    const int key = (m * 32 + T) * 8 + c;
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) return it->second;
    gf2t_fft_plan *p = new gf2t_fft_plan{m, T, schedule(m, true, T, c), schedule(m, false, T, c)};
    g_plans[key] = p;
    return p;
}
const gf2t_fft_plan *cached_plan(int m) { return cached_plan_T(m, tile_bits_for(m)); }

// Run the passes of one transform (or of nbatch transforms f_stride words
// apart, blockIdx.y selecting one); pk (may be null) packs the first
// pass's input.
int run_passes(const std::vector<PassDesc> &ps, u64 *df, hipStream_t st, const PackArg *pk,
               unsigned nbatch = 1, u64 f_stride = 0) {
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    std::call_once(lds_once, [] {
        lds_err = hipFuncSetAttribute((const void *)k_fft_pass,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)((4 * 256 + ((size_t)1 << FFT_TILE_BITS)) * sizeof(u64)));
    });
    GF2T_HIP(lds_err);
    if (nbatch == 0) return GF2T_OK;
    if (pk && pk->nlist && nbatch > (unsigned)pk->nlist) {
        gf2t::set_error("fft: %u transforms but %d pack sources", nbatch, pk->nlist);
        return GF2T_EINTERNAL;
    }
    PackArg none{};
    for (size_t i = 0; i < ps.size(); i++) {
        const PassDesc &pd = ps[i];
        const int T = pd.c + pd.w;
        const int ntab = std::max(1, std::min(4, (pd.m - 1 + 7) / 8));
        const unsigned blocks = (unsigned)(1ULL << (pd.m - T));
        const unsigned threads = threads_for_tile(T);
        const size_t lds = (4 * 256 + ((size_t)1 << T)) * sizeof(u64);
        const bool pack = i == 0 && pk;
        hipLaunchKernelGGL(k_fft_pass, dim3(blocks, nbatch), dim3(threads), lds, st, df, pd,
                           pack ? *pk : none, pack ? 1 : 0, ntab, f_stride);
    }
    GF2T_CHECK_LAUNCH("k_fft_pass");
    return GF2T_OK;
}

PackArg pack_one(const u64 *src, u64 bits, u32 xor0 = 0) {
    PackArg p{};
    p.src = (const u32 *)src;
    p.nchunks = (bits + 31) >> 5;
    p.lastmask = (bits & 31) ? (u32)((1ULL << (bits & 31)) - 1) : 0xFFFFFFFFu;
    p.xor0 = xor0;
    return p;
}

int fft_m_for(u64 nchunks_total) {
    int m = 1;
    while (((u64)1 << m) < nchunks_total) m++;
    return m;
}

}  // namespace

// gf2t_tune("fft_tile_<m>", T): tiles of 2^T elements (6..13) for 2^m-point
// transforms from now on; 0 restores the default.  Plans already made
// keep theirs.
int gf2t_fft_set_tile_bits(int m, int64_t T) {
    if (m < 1 || m > FFT_MAX_M || (T != 0 && (T < FFT_MIN_TILE_BITS || T > FFT_TILE_BITS))) {
        gf2t::set_error("gf2t_tune: fft_tile_<m> needs 1 <= m <= %d and T = 0 or %d..%d",
                        FFT_MAX_M, FFT_MIN_TILE_BITS, FFT_TILE_BITS);
        return GF2T_EINVAL;
    }
    g_tile_bits[m].store((int)T);
    return GF2T_OK;
}
int gf2t_fft_tile_bits(int m) { return (m >= 1 && m <= FFT_MAX_M) ? tile_bits_for(m) : 0; }
int gf2t_fft_set_ept(int64_t ept) {
    if (ept != 2 && ept != 4 && ept != 8 && ept != 16) {
        gf2t::set_error("gf2t_tune: fft_ept needs 2, 4, 8 or 16");
        return GF2T_EINVAL;
    }
    g_fft_ept.store((int)ept);
    return GF2T_OK;
}
int gf2t_fft_ept() { return g_fft_ept.load(); }
int gf2t_fft_set_low_bits(int64_t c) {
    if (c < 4 || c > 6) {
        gf2t::set_error("gf2t_tune: fft_low_bits needs 4, 5 or 6");
        return GF2T_EINVAL;
    }
    g_fft_low.store((int)c);
    return GF2T_OK;
}
int gf2t_fft_low_bits() { return g_fft_low.load(); }

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
    const int T = std::max(FFT_MIN_TILE_BITS, std::min(tile_bits_for(m), FFT_TILE_BITS));
    const int c = g_fft_low.load(std::memory_order_relaxed);
    *plan = new gf2t_fft_plan{m, T, schedule(m, true, T, c), schedule(m, false, T, c)};
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
    return run_passes(plan->fwd, df, (hipStream_t)stream, nullptr);
}

extern "C" int gf2t_fft_inv(const gf2t_fft_plan *plan, uint64_t *df, void *stream) {
    if (!plan || !df) return GF2T_EINVAL;
    int rc = ensure_basis();
    if (rc) return rc;
    return run_passes(plan->inv, df, (hipStream_t)stream, nullptr);
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
// p[0, n) = 0 with a kernel (no memset node when the caller captures a graph)
__global__ void k_zero_words(u64 *__restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}
int zero_words(u64 *p, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)),
                       dim3(256), 0, st, p, n);
    GF2T_CHECK_LAUNCH("k_zero_words");
    return GF2T_OK;
}

// ---------------------------------------------------- wrapped products
// A product whose N = ca + cb - 1 chunk coefficients lie just above a
// power of two, 2^mp < N <= 2^mp + 2^mp/4, runs on 2^mp-point transforms
// instead of 2^(mp+1), plus a product at most half that size.  The 2^mp
// points are the span of the first mp Cantor basis elements, whose
// vanishing polynomial is
//     V(X) = sum over k with bits(k) inside bits(mp) of X^(2^k)
// (V_{i+1} = V_i^2 + V_i, since V_i(b_i) = 1 for the Cantor basis; the
// binomial coefficients mod 2 follow Lucas).  So pointwise products and
// the inverse give C mod V.  The e = N - 2^mp top coefficients
// C_high = C div X^(2^mp) come from the product of the top e chunks of
// each operand (pairs i + j >= 2^mp only involve those), a 2^ms-point
// product.  V's next term below X^(2^mp) is at most X^(2^(mp-1)), and
// deg C_high < e <= 2^(mp-2), so C div V = C_high and
//     C = (C mod V) + C_high * (V - X^(2^mp)) + C_high * X^(2^mp):
// C_high is XORed into the low coefficients at each offset 2^k, k a
// proper bit-subset of mp, and forms the coefficients from 2^mp up.
struct Wrap {
    bool on = false;
    int mp = 0, ms = 0;
    u64 e = 0;
};
struct WrapArg {
    u64 n;          // 2^mp
    u64 e;          // top coefficients
    int noff;
    u32 off[15];    // 2^k for the proper bit-subsets k of mp
};
Wrap wrap_for(u64 ca, u64 cb) {
    Wrap w;
    const u64 N = ca + cb - 1;
    const int m = fft_m_for(N);
    if (m < 12 || m - 1 > FFT_MAX_M) return w;
    const int mp = m - 1;
    const u64 n = (u64)1 << mp;
    if (N <= n) return w;
    const u64 e = N - n;
    if (e > n / 4 || ca > n || cb > n || e >= ca || e >= cb) return w;
    w.on = true;
    w.mp = mp;
    w.e = e;
    w.ms = fft_m_for(2 * e - 1);
    return w;
}
WrapArg wrap_arg(const Wrap &w) {
    WrapArg a{};
    a.n = (u64)1 << w.mp;
    a.e = w.e;
    for (int k = 0; k < w.mp; k++)
        if ((k & w.mp) == k) a.off[a.noff++] = 1u << k;
    return a;
}

// chunk coefficient k < n + e of the wrapped product: R = C mod V (2^mp
// coefficients), H = C_high
GF2T_DEV u64 wrap_chunk(const u64 *__restrict__ R, const u64 *__restrict__ H, const WrapArg &w,
                        u64 k) {
    if (k >= w.n) return H[k - w.n];
    u64 v = R[k];
    for (int j = 0; j < w.noff; j++) {
        const u64 o = w.off[j];
        if (k >= o && k - o < w.e) v ^= H[k - o];
    }
    return v;
}

// overlap-add of a wrapped product (gather form, as k_overlap_add);
// blockIdx.y selects a batch element: R + y*r_stride, H + y*h_stride,
// out + y*o_stride
__global__ void k_overlap_add_wrap(const u64 *__restrict__ R, u64 r_stride,
                                   const u64 *__restrict__ H, u64 h_stride, const WrapArg w,
                                   size_t nchunks, u64 *__restrict__ out, size_t owords,
                                   u64 o_stride) {
    size_t wd = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (wd >= owords) return;
    R += (u64)blockIdx.y * r_stride;
    H += (u64)blockIdx.y * h_stride;
    out += (u64)blockIdx.y * o_stride;
    size_t k = 2 * wd;
    u64 v = 0;
    if (k < nchunks) v ^= wrap_chunk(R, H, w, k);
    if (k + 1 < nchunks) v ^= wrap_chunk(R, H, w, k + 1) << 32;
    if (wd && k - 1 < nchunks) v ^= wrap_chunk(R, H, w, k - 1) >> 32;
    out[wd] = v;
}

// pack source of chunks [first, first + count) of an operand of nch chunks
// whose last chunk is masked to lastmask
PackArg pack_range(const u32 *src, u64 first, u64 count, u64 nch, u32 lastmask) {
    PackArg p{};
    p.src = src + first;
    p.nchunks = count;
    p.lastmask = first + count == nch ? lastmask : 0xFFFFFFFFu;
    return p;
}

int mul_core(const u64 *da, u64 abits, const u64 *db, u64 bbits, u64 *dprod, void *dwork,
             hipStream_t st) {
    u64 ca = (abits + 31) >> 5, cb = (bbits + 31) >> 5;
    size_t ow = gf2t_mul_out_words(abits, bbits);
    if (!ca || !cb) return zero_words(dprod, ow, st);
    int m = fft_m_for(ca + cb - 1);
    if (m > FFT_MAX_M) {
        gf2t::set_error("gf2t_mul: product needs 2^%d points (max 2^%d)", m, FFT_MAX_M);
        return GF2T_ERANGE;
    }
    int rc = ensure_basis();
    if (rc) return rc;
    const Wrap W = wrap_for(ca, cb);
    if (W.on) {
        // both operands at 2^mp points, their top e chunks at 2^ms points
        // (2^mp + 2^ms <= 2^m words each, inside the workspace of an
        // unwrapped product)
        const gf2t_fft_plan *pl = cached_plan(W.mp), *ps = cached_plan(W.ms);
        const size_t n = (size_t)1 << W.mp, ns = (size_t)1 << W.ms;
        u64 *FA = (u64 *)dwork, *FB = FA + n, *TA = FB + n, *TB = TA + ns;
        const PackArg pa = pack_one(da, abits), pb = pack_one(db, bbits);
        PackArg two{};
        two.nlist = 2;
        two.lsrc[0] = pa.src; two.lnch[0] = pa.nchunks; two.lmask[0] = pa.lastmask;
        two.lsrc[1] = pb.src; two.lnch[1] = pb.nchunks; two.lmask[1] = pb.lastmask;
        rc = run_passes(pl->fwd, FA, st, &two, 2, n);
        if (rc) return rc;
        const PackArg ta = pack_range(pa.src, ca - W.e, W.e, ca, pa.lastmask);
        const PackArg tb = pack_range(pb.src, cb - W.e, W.e, cb, pb.lastmask);
        PackArg top{};
        top.nlist = 2;
        top.lsrc[0] = ta.src; top.lnch[0] = ta.nchunks; top.lmask[0] = ta.lastmask;
        top.lsrc[1] = tb.src; top.lnch[1] = tb.nchunks; top.lmask[1] = tb.lastmask;
        rc = run_passes(ps->fwd, TA, st, &top, 2, ns);
        if (rc) return rc;
        hipLaunchKernelGGL(k_pointwise, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)),
                           dim3(256), 0, st, FA, FB, n);
        hipLaunchKernelGGL(k_pointwise, dim3((unsigned)std::min<size_t>((ns + 255) / 256, 8192)),
                           dim3(256), 0, st, TA, TB, ns);
        rc = run_passes(pl->inv, FA, st, nullptr);
        if (rc) return rc;
        rc = run_passes(ps->inv, TA, st, nullptr);
        if (rc) return rc;
        // C_high[u] = coefficient e - 1 + u of the top product
        hipLaunchKernelGGL(k_overlap_add_wrap, dim3((unsigned)((ow + 255) / 256)), dim3(256), 0,
                           st, FA, (u64)0, TA + (W.e - 1), (u64)0, wrap_arg(W),
                           (size_t)(ca + cb - 1), dprod, ow, (u64)0);
        GF2T_CHECK_LAUNCH("k_overlap_add_wrap");
        return GF2T_OK;
    }
    const gf2t_fft_plan *plan = cached_plan(m);
    size_t n = (size_t)1 << m;
    u64 *FA = (u64 *)dwork, *FB = FA + n;
    // both operands in one launch chain: FA and FB are n words apart
    const PackArg pa = pack_one(da, abits), pb = pack_one(db, bbits);
    PackArg two{};
    two.nlist = 2;
    two.lsrc[0] = pa.src; two.lnch[0] = pa.nchunks; two.lmask[0] = pa.lastmask;
    two.lsrc[1] = pb.src; two.lnch[1] = pb.nchunks; two.lmask[1] = pb.lastmask;
    rc = run_passes(plan->fwd, FA, st, &two, 2, n);
    if (rc) return rc;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_pointwise, dim3(blocks), dim3(256), 0, st, FA, FB, n);
    rc = run_passes(plan->inv, FA, st, nullptr);
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
    if (!(r >= 3 && s >= 1 && 2 * s < r)) {
        gf2t::set_error("gf2t_mulmod: need 1 <= s, 2s < r");
        return GF2T_EINVAL;
    }
    u64 ca = (abits + 31) >> 5, cb = (bbits + 31) >> 5;
    if (!ca || !cb)   // a zero operand: the workspace may be empty, dout is 0
        return zero_words(dout, words_of(r), (hipStream_t)stream);
    size_t n = (size_t)1 << fft_m_for(ca + cb - 1);
    u64 *prod = (u64 *)dwork + 2 * n;
    int rc = mul_core(da, abits, db, bbits, prod, dwork, (hipStream_t)stream);
    if (rc) return rc;
    return gf2t_fold(prod, abits + bbits - 1, dout, r, s, stream);
}

// ------------------------------------------------ modmul without allocation
// gf2t_mulmod_noalloc: gf2t_mulmod with its workspace in a library-owned
// device array instead of a caller's buffer, for callers that cannot
// allocate.  The array holds the two 2^22-point transforms and the
// product (up to 2^27 bits): 80 MiB, which every GPU the library runs on
// keeps until the process ends.  The pack pass writes every transform
// element and the overlap-add every product word, so nothing in it needs
// clearing between calls.  Calls that overlap on one GPU share the array
// and race.
namespace {
constexpr int NA_MAX_M = 22;
constexpr size_t NA_FFT_WORDS = (size_t)1 << NA_MAX_M;
constexpr size_t NA_PROD_WORDS = ((size_t)1 << 21) + 2;   // gf2t_mul_out_words(a, b), a+b <= 2^27+64
constexpr size_t NA_WORDS = 2 * NA_FFT_WORDS + NA_PROD_WORDS + 30;
constexpr int NA_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) u64 scratch[NA_WORDS];

namespace {
std::mutex g_scratch_mu;
// device address of `scratch` on each device, from hipGetSymbolAddress
// (the name itself is only a host-side stand-in)
u64 *g_scratch_ptr[NA_MAX_DEVICES] = {};

}  // namespace

// Once per device, with that device current (gf2t_init calls it; the
// entry point calls it on its first use on a device otherwise): record
// the scratch array's device address, upload the Cantor basis and build
// the transform plans the entry point can need, so that its calls
// allocate nothing.
int gf2t_noalloc_setup() {
    int dev = 0;
    GF2T_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        gf2t::set_error("gf2t_noalloc_setup: device %d beyond %d", dev, NA_MAX_DEVICES);
        return GF2T_ERANGE;
    }
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void *p = nullptr;
            GF2T_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch_ptr[dev] = (u64 *)p;
        }
    }
    for (int m = 1; m <= NA_MAX_M; m++) cached_plan(m);
    return ensure_basis();
}

extern "C" int gf2t_mulmod_noalloc(const uint64_t *da, uint64_t abits, const uint64_t *db,
                                   uint64_t bbits, uint64_t r, uint64_t s, uint64_t *dout,
                                   void *stream) {
    if (!da || !db || !dout) {
        gf2t::set_error("gf2t_mulmod_noalloc: null pointer");
        return GF2T_EINVAL;
    }
    if (!(r >= 3 && s >= 1 && 2 * s < r) || abits + bbits > 2 * r) {
        gf2t::set_error("gf2t_mulmod_noalloc: need 1 <= s, 2s < r and a product of at most "
                        "2r-1 bits (r=%llu s=%llu, %llu+%llu bits)", (unsigned long long)r,
                        (unsigned long long)s, (unsigned long long)abits,
                        (unsigned long long)bbits);
        return GF2T_EINVAL;
    }
    const u64 ca = (abits + 31) >> 5, cb = (bbits + 31) >> 5;
    if (ca && cb && ca + cb - 1 > NA_FFT_WORDS) {
        gf2t::set_error("gf2t_mulmod_noalloc: a %llu x %llu-bit product needs more than the "
                        "2^%d transform points its scratch holds (at most %llu 32-bit chunks "
                        "in both operands together); use gf2t_mulmod with a workspace",
                        (unsigned long long)abits, (unsigned long long)bbits, NA_MAX_M,
                        (unsigned long long)(NA_FFT_WORDS + 1));
        return GF2T_ERANGE;
    }
    int dev = 0;
    GF2T_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        gf2t::set_error("gf2t_mulmod_noalloc: device %d beyond %d", dev, NA_MAX_DEVICES);
        return GF2T_ERANGE;
    }
    u64 *ptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        ptr = g_scratch_ptr[dev];
    }
    if (!ptr) {
        int rc = gf2t_noalloc_setup();
        if (rc) return rc;
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        ptr = g_scratch_ptr[dev];
    }
    hipStream_t st = (hipStream_t)stream;
    if (!ca || !cb) return zero_words(dout, words_of(r), st);   // a zero operand
    const size_t n = (size_t)1 << fft_m_for(ca + cb - 1);
    u64 *prod = ptr + 2 * n;
    int rc = mul_core(da, abits, db, bbits, prod, ptr, st);
    if (rc) return rc;
    return gf2t_fold(prod, abits + bbits - 1, dout, r, s, stream);
}

// ------------------------------------------------- variable-size batch
// gf2t_mulmod_batch: products of different sizes, each folded mod
// x^r + x^s_j + 1, in one set of launches: the i-th forward (inverse) pass
// of every product that has one shares a grouped launch, and the
// pointwise products, overlap-adds and folds of all pairs are one launch
// each.  Each product keeps its own transform size and tile.
namespace {
struct PairSeg {       // one pair's part of a flat grouped launch over elements
    u64 *fa;
    const u64 *fb;
    u64 *prod;
    u64 *out;
    u64 n, nch, ow, pw, L, s, start;   // start: the pair's first element in the launch
};

GF2T_DEV int find_seg(const PairSeg *__restrict__ sg, int np, u64 i) {
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sg[mid].start <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// phase 0: FA *= FB pointwise; 1: overlap-add FA into prod; 2: fold prod
// into out.  total = sum of the pairs' element counts for the phase.
__global__ void k_pairs_grouped(const PairSeg *__restrict__ sg, int np, u64 total, int phase,
                                u64 r) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < total; i += stride) {
        const PairSeg &S = sg[find_seg(sg, np, i)];
        const u64 k = i - S.start;
        if (phase == 0) {
            S.fa[k] = gf64_mul(S.fa[k], S.fb[k]);
        } else if (phase == 1) {
            const u64 c = 2 * k;
            u64 v = 0;
            if (c < S.nch) v ^= S.fa[c];
            if (c + 1 < S.nch) v ^= S.fa[c + 1] << 32;
            if (k && c - 1 < S.nch) v ^= S.fa[c - 1] >> 32;
            S.prod[k] = v;
        } else {
            S.out[k] = fold_word(S.prod, S.pw, S.L, r, S.s, k);
        }
    }
}

struct BatchLayout {
    std::vector<size_t> fa, prod;   // word offsets of each pair's transforms and product
    std::vector<int> m;
    size_t desc_off = 0, words = 0;
};

BatchLayout batch_layout(int n, const uint64_t *abits, const uint64_t *bbits, size_t ndesc_bytes) {
    BatchLayout L;
    size_t off = 0;
    for (int j = 0; j < n; j++) {
        u64 ca = (abits[j] + 31) >> 5, cb = (bbits[j] + 31) >> 5;
        int m = (ca && cb) ? fft_m_for(ca + cb - 1) : 0;
        L.m.push_back(m);
        L.fa.push_back(off);
        off += m ? 2 * ((size_t)1 << m) : 0;
        L.prod.push_back(off);
        off += (gf2t_mul_out_words(abits[j], bbits[j]) + 31) & ~(size_t)31;
    }
    L.desc_off = off;
    L.words = off + (ndesc_bytes + 7) / 8 + 32;
    return L;
}

// room for the descriptors whatever tiles the plans use: a pass holds at
// least one op, so a transform has at most m(m+1)/2 passes each way
size_t batch_desc_bytes(int n, const std::vector<int> &ms) {
    size_t steps = 0;
    for (int m : ms)
        if (m) steps += 2 * (size_t)(m * (m + 1) / 2);
    return steps * sizeof(GroupPass) + (size_t)n * 3 * sizeof(PairSeg) + 256;
}
}  // namespace

extern "C" size_t gf2t_mulmod_batch_workspace_bytes(int n, const uint64_t *abits,
                                                   const uint64_t *bbits) {
    if (n <= 0 || !abits || !bbits) return 0;
    BatchLayout L0 = batch_layout(n, abits, bbits, 0);
    return batch_layout(n, abits, bbits, batch_desc_bytes(n, L0.m)).words * 8;
}

extern "C" int gf2t_mulmod_batch(int n, const uint64_t *const *da, const uint64_t *abits,
                                 const uint64_t *const *db, const uint64_t *bbits, uint64_t r,
                                 const uint64_t *s, uint64_t *const *dout, void *dwork,
                                 void *stream) {
    if (n < 0 || (n > 0 && (!da || !abits || !db || !bbits || !s || !dout || !dwork))) {
        gf2t::set_error("gf2t_mulmod_batch: bad arguments");
        return GF2T_EINVAL;
    }
    if (n == 0) return GF2T_OK;
    for (int j = 0; j < n; j++) {
        if (!da[j] || !db[j] || !dout[j] || !(r >= 3 && s[j] >= 1 && 2 * s[j] < r) ||
            abits[j] + bbits[j] > 2 * r) {
            gf2t::set_error("gf2t_mulmod_batch: pair %d: bad pointers, s, or a product over "
                            "2r-1 bits", j);
            return GF2T_EINVAL;
        }
        u64 ca = (abits[j] + 31) >> 5, cb = (bbits[j] + 31) >> 5;
        if (ca && cb && fft_m_for(ca + cb - 1) > FFT_MAX_M) {
            gf2t::set_error("gf2t_mulmod_batch: pair %d needs more than 2^%d points", j,
                            FFT_MAX_M);
            return GF2T_ERANGE;
        }
    }
    int rc = ensure_basis();
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    BatchLayout L0 = batch_layout(n, abits, bbits, 0);
    BatchLayout L = batch_layout(n, abits, bbits, batch_desc_bytes(n, L0.m));
    u64 *W = (u64 *)dwork;
    // the steps of the grouped launches: forward passes, then inverse
    std::vector<const gf2t_fft_plan *> plan(n, nullptr);
    size_t maxf = 0, maxi = 0;
    for (int j = 0; j < n; j++)
        if (L.m[j]) {
            plan[j] = cached_plan(L.m[j]);
            maxf = std::max(maxf, plan[j]->fwd.size());
            maxi = std::max(maxi, plan[j]->inv.size());
        }
    std::vector<GroupPass> desc;
    struct Step { size_t first; int count; u32 blocks; int T; };
    std::vector<Step> steps;
    for (int dir = 0; dir < 2; dir++) {
        const size_t ns = dir ? maxi : maxf;
        for (size_t i = 0; i < ns; i++) {
            Step S{desc.size(), 0, 0, 0};
            for (int j = 0; j < n; j++) {
                if (!plan[j]) continue;
                const std::vector<PassDesc> &ps = dir ? plan[j]->inv : plan[j]->fwd;
                if (i >= ps.size()) continue;
                GroupPass g{};
                g.pd = ps[i];
                const size_t nn = (size_t)1 << L.m[j];
                g.f = W + L.fa[j];
                g.f_stride = nn;
                g.ntrans = dir ? 1 : 2;
                g.pack = (!dir && i == 0) ? 1 : 0;
                g.src[0] = (const u32 *)da[j];
                g.src[1] = (const u32 *)db[j];
                g.nch[0] = (abits[j] + 31) >> 5;
                g.nch[1] = (bbits[j] + 31) >> 5;
                g.mask[0] = (abits[j] & 31) ? (u32)((1ULL << (abits[j] & 31)) - 1) : 0xFFFFFFFFu;
                g.mask[1] = (bbits[j] & 31) ? (u32)((1ULL << (bbits[j] & 31)) - 1) : 0xFFFFFFFFu;
                g.ntab = std::max(1, std::min(4, (L.m[j] - 1 + 7) / 8));
                const int T = g.pd.c + g.pd.w;
                g.nblocks = (u32)(1ULL << (g.pd.m - T));
                g.block0 = S.blocks;
                S.blocks += g.nblocks * g.ntrans;
                S.T = std::max(S.T, T);
                S.count++;
                desc.push_back(g);
            }
            steps.push_back(S);
        }
    }
    std::vector<PairSeg> seg[3];
    u64 total[3] = {0, 0, 0};
    int nz = 0;
    for (int j = 0; j < n; j++) {
        const size_t nw = words_of(r);
        if (!L.m[j]) {   // a zero operand: the residue is 0
            rc = zero_words(dout[j], nw, st);
            if (rc) return rc;
            nz++;
            continue;
        }
        const u64 nn = 1ULL << L.m[j], ca = (abits[j] + 31) >> 5, cb = (bbits[j] + 31) >> 5;
        const u64 ow = gf2t_mul_out_words(abits[j], bbits[j]);
        PairSeg p{};
        p.fa = W + L.fa[j];
        p.fb = p.fa + nn;
        p.prod = W + L.prod[j];
        p.out = dout[j];
        p.n = nn;
        p.nch = ca + cb - 1;
        p.ow = ow;
        p.pw = ow;
        p.L = abits[j] + bbits[j] - 1;
        p.s = s[j];
        const u64 cnt[3] = {nn, ow, (u64)nw};
        for (int ph = 0; ph < 3; ph++) {
            p.start = total[ph];
            total[ph] += cnt[ph];
            seg[ph].push_back(p);
        }
    }
    const int np = n - nz;
    if (np == 0) return GF2T_OK;
    // descriptors: the pass entries, then the three segment tables
    std::vector<unsigned char> host(desc.size() * sizeof(GroupPass) + 3 * np * sizeof(PairSeg));
    std::memcpy(host.data(), desc.data(), desc.size() * sizeof(GroupPass));
    for (int ph = 0; ph < 3; ph++)
        std::memcpy(host.data() + desc.size() * sizeof(GroupPass) + ph * np * sizeof(PairSeg),
                    seg[ph].data(), np * sizeof(PairSeg));
    unsigned char *dd = (unsigned char *)(W + L.desc_off);
    dd = (unsigned char *)(((uintptr_t)dd + 255) & ~(uintptr_t)255);
    GF2T_HIP(hipMemcpyAsync(dd, host.data(), host.size(), hipMemcpyHostToDevice, st));
    GF2T_HIP(hipStreamSynchronize(st));   // `host` is a pageable local buffer
    const GroupPass *dgp = (const GroupPass *)dd;
    const PairSeg *dseg = (const PairSeg *)(dd + desc.size() * sizeof(GroupPass));
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    std::call_once(lds_once, [] {
        lds_err = hipFuncSetAttribute((const void *)k_fft_pass_grouped,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)((4 * 256 + ((size_t)1 << FFT_TILE_BITS)) * sizeof(u64)));
    });
    GF2T_HIP(lds_err);
    auto grouped_phase = [&](int ph) {
        const u64 tot = total[ph];
        unsigned blocks = (unsigned)std::min<u64>((tot + 255) / 256, 65536);
        hipLaunchKernelGGL(k_pairs_grouped, dim3(blocks), dim3(256), 0, st, dseg + ph * np, np,
                           tot, ph, (u64)r);
    };
    for (size_t k = 0; k < steps.size(); k++) {
        if (k == maxf) grouped_phase(0);   // between the forward and inverse passes
        const Step &S = steps[k];
        const size_t lds = (4 * 256 + ((size_t)1 << S.T)) * sizeof(u64);
        const unsigned threads =
            threads_for_tile(S.T);
        hipLaunchKernelGGL(k_fft_pass_grouped, dim3(S.blocks), dim3(threads), lds, st,
                           dgp + S.first, S.count);
    }
    if (steps.size() == maxf) grouped_phase(0);
    grouped_phase(1);
    grouped_phase(2);
    GF2T_CHECK_LAUNCH("gf2t_mulmod_batch");
    return GF2T_OK;
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
    const Wrap W = wrap_for(ca, ca);
    if (W.on) {
        // the wrapped product (see wrap_for): 2^mp-point transforms of both
        // operands and 2^ms-point transforms of their top e chunks, laid
        // out inside the 2 N n words an unwrapped product uses
        const gf2t_fft_plan *pl = cached_plan(W.mp), *ps = cached_plan(W.ms);
        const size_t np = (size_t)1 << W.mp, ns = (size_t)1 << W.ms;
        u64 *GA = (u64 *)dwork, *GB = GA + np * n, *TA = GB + np * n, *TB = TA + ns * n;
        PackArg pa{};
        pa.src = (const u32 *)da;
        pa.stride = 2 * apitch;
        pa.nchunks = ca;
        pa.lastmask = lm;
        rc = run_passes(pl->fwd, GA, st, &pa, (unsigned)n, np);
        if (rc) return rc;
        PackArg pb = pa;
        pb.src = (const u32 *)db;
        pb.stride = 2 * bpitch;
        pb.xor0 = b_add_x ? 2u : 0u;
        rc = run_passes(pl->fwd, GB, st, &pb, (unsigned)n, np);
        if (rc) return rc;
        // top chunks [ca - e, ca); chunk 0 (where b_add_x adds x) is not among them
        PackArg ta = pa, tb = pb;
        ta.src += ca - W.e;
        tb.src += ca - W.e;
        ta.nchunks = tb.nchunks = W.e;
        tb.xor0 = 0;
        rc = run_passes(ps->fwd, TA, st, &ta, (unsigned)n, ns);
        if (rc) return rc;
        rc = run_passes(ps->fwd, TB, st, &tb, (unsigned)n, ns);
        if (rc) return rc;
        size_t tot = np * (size_t)n, tots = ns * (size_t)n;
        hipLaunchKernelGGL(k_pointwise, dim3((unsigned)std::min<size_t>((tot + 255) / 256, 65536)),
                           dim3(256), 0, st, GA, GB, tot);
        hipLaunchKernelGGL(k_pointwise, dim3((unsigned)std::min<size_t>((tots + 255) / 256, 65536)),
                           dim3(256), 0, st, TA, TB, tots);
        rc = run_passes(pl->inv, GA, st, nullptr, (unsigned)n, np);
        if (rc) return rc;
        rc = run_passes(ps->inv, TA, st, nullptr, (unsigned)n, ns);
        if (rc) return rc;
        size_t ow = 2 * nw;
        hipLaunchKernelGGL(k_overlap_add_wrap, dim3((unsigned)((ow + 255) / 256), n), dim3(256), 0,
                           st, GA, (u64)np, TA + (W.e - 1), (u64)ns, wrap_arg(W),
                           (size_t)(2 * ca - 1), prod, ow, (u64)ppitch);
        hipLaunchKernelGGL(k_fold_slab, dim3((unsigned)((nw + 255) / 256), n), dim3(256), 0, st,
                           prod, ow, (u64)ppitch, 2 * r - 1, dout, (u64)opitch, nw, r, ds);
        GF2T_CHECK_LAUNCH("gf2t_mulmod_slab");
        return GF2T_OK;
    }
    PackArg pa{};
    pa.src = (const u32 *)da;
    pa.stride = 2 * apitch;
    pa.nchunks = ca;
    pa.lastmask = lm;
    rc = run_passes(plan->fwd, FA, st, &pa, (unsigned)n, N);
    if (rc) return rc;
    PackArg pb = pa;
    pb.src = (const u32 *)db;
    pb.stride = 2 * bpitch;
    pb.xor0 = b_add_x ? 2u : 0u;
    rc = run_passes(plan->fwd, FB, st, &pb, (unsigned)n, N);
    if (rc) return rc;
    size_t tot = N * (size_t)n;
    unsigned blocks = (unsigned)std::min<size_t>((tot + 255) / 256, 65536);
    hipLaunchKernelGGL(k_pointwise, dim3(blocks), dim3(256), 0, st, FA, FB, tot);
    rc = run_passes(plan->inv, FA, st, nullptr, (unsigned)n, N);
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

// A pool of contexts, each with its own non-blocking stream, device
// buffer and pinned host staging buffer, so products requested by
// different host threads (the gcds of different s values) run
// concurrently on the device.  Operands are copied into the pinned
// buffer and go to the device in one transfer; results come back the
// same way.
constexpr int HM_CONTEXTS = 16;
// products whose operands and results together take at most this many
// words go through pinned host memory directly (gf2t_tune "hm_zc_words")
std::atomic<long long> g_hm_zc_words{65536};
struct HostMulCtx {
    hipStream_t st = nullptr;
    u64 *dbuf = nullptr;
    size_t dcap = 0;   // words
    u64 *hbuf = nullptr;
    size_t hcap = 0;   // words
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

int hm_reserve(HostMulCtx &c, size_t words, size_t host_words) {
    if (!c.st) GF2T_HIP(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    if (words > c.dcap) {
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
    }
    if (host_words > c.hcap) {
        if (c.hbuf) GF2T_HIP(hipHostFree(c.hbuf));
        c.hbuf = nullptr;
        c.hcap = 0;
        size_t want = host_words + host_words / 4;
        if (hipHostMalloc((void **)&c.hbuf, want * 8, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            gf2t::set_error("gf2t host product service: cannot allocate %zu pinned bytes",
                            want * 8);
            return GF2T_ENOMEM;
        }
        c.hcap = want;
    }
    return GF2T_OK;
}

u32 last_mask(u64 bits) { return (bits & 31) ? (u32)((1ULL << (bits & 31)) - 1) : 0xFFFFFFFFu; }
}  // namespace

int gf2t_hm_set_zc_words(int64_t words) {
    if (words < 0) {
        gf2t::set_error("gf2t_tune: hm_zc_words must be >= 0");
        return GF2T_EINVAL;
    }
    g_hm_zc_words.store(words);
    return GF2T_OK;
}
int64_t gf2t_hm_zc_words() { return g_hm_zc_words.load(); }

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
    const size_t aw = words_of(abits), bw = words_of(bbits);
    const size_t ww = (gf2t_mul_workspace_bytes(abits, bbits) + 7) / 8;
    int rc = hm_reserve(hc, aw + bw + ow + ww + 64, std::max(aw + bw, ow));
    if (rc) return rc;
    u64 *da = hc.dbuf, *db = da + aw, *dout = db + bw, *dw = dout + ow;
    dw = (u64 *)(((uintptr_t)dw + 255) & ~(uintptr_t)255);
    hipStream_t st = hc.st;
    std::memcpy(hc.hbuf, ha, aw * 8);
    std::memcpy(hc.hbuf + aw, hb, bw * 8);
    GF2T_HIP(hipMemcpyAsync(da, hc.hbuf, (aw + bw) * 8, hipMemcpyHostToDevice, st));
    rc = mul_core(da, abits, db, bbits, dout, dw, st);
    if (rc) return rc;
    GF2T_HIP(hipMemcpyAsync(hc.hbuf, dout, ow * 8, hipMemcpyDeviceToHost, st));
    GF2T_HIP(hipStreamSynchronize(st));
    std::memcpy(hout, hc.hbuf, ow * 8);
    return GF2T_OK;
}

namespace {
// Pair p of a batched matrix-vector product: its two transforms (at F + (4
// + 2p) n and F + (5 + 2p) n) become (m0 a + m1 b, m2 a + m3 b), the
// matrix transforms at F .. F + 3n left as they are.
__global__ void k_mat2n_pointwise(u64 *__restrict__ F, size_t n, int npairs) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const u64 m0 = F[i], m1 = F[n + i], m2 = F[2 * n + i], m3 = F[3 * n + i];
        for (int p = 0; p < npairs; p++) {
            u64 *fa = F + (4 + 2 * (size_t)p) * n, *fb = fa + n;
            const u64 a = fa[i], b = fb[i];
            fa[i] = gf64_mul(m0, a) ^ gf64_mul(m1, b);
            fb[i] = gf64_mul(m2, a) ^ gf64_mul(m3, b);
        }
    }
}
}  // namespace

extern "C" int gf2t_mat2n_host(const uint64_t *const hm[4], const uint64_t mbits[4], int npairs,
                               const uint64_t *const *hv, const uint64_t *vbits,
                               uint64_t *const *hout, size_t outw) {
    if (!hm || !mbits || !hv || !vbits || !hout || npairs < 1 || npairs > (PACK_LIST - 4) / 2) {
        gf2t::set_error("gf2t_mat2n_host: bad arguments (1 <= npairs <= %d)", (PACK_LIST - 4) / 2);
        return GF2T_EINVAL;
    }
    const int nv = 2 * npairs;
    u64 mb = 0, vb = 0;
    for (int i = 0; i < 4; i++) {
        if (!hm[i] && mbits[i]) { gf2t::set_error("gf2t_mat2n_host: null entry"); return GF2T_EINVAL; }
        mb = std::max<u64>(mb, mbits[i]);
    }
    for (int i = 0; i < nv; i++) {
        if ((!hv[i] && vbits[i]) || !hout[i]) {
            gf2t::set_error("gf2t_mat2n_host: null operand");
            return GF2T_EINVAL;
        }
        vb = std::max<u64>(vb, vbits[i]);
    }
    if (outw < gf2t_mul_out_words(mb, vb)) {
        gf2t::set_error("gf2t_mat2n_host: outw %zu below %zu", outw, gf2t_mul_out_words(mb, vb));
        return GF2T_EINVAL;
    }
    if (!mb || !vb) {
        for (int i = 0; i < nv; i++) std::memset(hout[i], 0, outw * 8);
        return GF2T_OK;
    }
    u64 cm = (mb + 31) >> 5, cv = (vb + 31) >> 5;
    int m = fft_m_for(cm + cv - 1);
    if (m > FFT_MAX_M) { gf2t::set_error("gf2t_mat2n_host: product too large"); return GF2T_ERANGE; }
    int rc = ensure_basis();
    if (rc) return rc;
    HmLease lease;
    HostMulCtx &hc = *lease.c;
    const gf2t_fft_plan *plan = cached_plan(m);
    const int nt = 4 + nv;                    // transforms: matrix, then the vectors
    size_t n = (size_t)1 << m, ow = gf2t_mul_out_words(mb, vb);
    const u64 *src[PACK_LIST];
    u64 bits[PACK_LIST];
    size_t in_w[PACK_LIST], stage = 0;
    for (int i = 0; i < nt; i++) {
        src[i] = i < 4 ? hm[i] : hv[i - 4];
        bits[i] = i < 4 ? mbits[i] : vbits[i - 4];
        in_w[i] = words_of(bits[i]);
        stage += (in_w[i] + 3) & ~(size_t)3;
    }
    rc = hm_reserve(hc, nt * n + nv * ow + stage + 64, std::max(stage, nv * ow));
    if (rc) return rc;
    hipStream_t st = hc.st;
    u64 *F = hc.dbuf;                         // nt transforms, n words apart
    u64 *dout = F + nt * n, *dstage = dout + nv * ow;
    // the operands, 32-byte aligned, in one pinned buffer and one copy
    PackArg pk{};
    pk.nlist = nt;
    size_t off = 0;
    for (int i = 0; i < nt; i++) {
        if (in_w[i]) std::memcpy(hc.hbuf + off, src[i], in_w[i] * 8);
        pk.lsrc[i] = (const u32 *)(dstage + off);
        pk.lnch[i] = (bits[i] + 31) >> 5;     // 0 for a zero operand: its transform is all zero
        pk.lmask[i] = last_mask(bits[i]);
        off += (in_w[i] + 3) & ~(size_t)3;
    }
    // small products read their operands from the pinned staging buffer
    // and write their results back into it (the GPU reaches pinned host
    // memory directly), instead of a copy each way
    const bool zc = off + nv * ow <= (size_t)g_hm_zc_words.load(std::memory_order_relaxed);
    if (zc) {
        size_t o2 = 0;
        for (int i = 0; i < nt; i++) {
            pk.lsrc[i] = (const u32 *)(hc.hbuf + o2);
            o2 += (in_w[i] + 3) & ~(size_t)3;
        }
        dout = hc.hbuf;
    } else if (off) {
        GF2T_HIP(hipMemcpyAsync(dstage, hc.hbuf, off * 8, hipMemcpyHostToDevice, st));
    }
    rc = run_passes(plan->fwd, F, st, &pk, (unsigned)nt, n);
    if (rc) return rc;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_mat2n_pointwise, dim3(blocks), dim3(256), 0, st, F, n, npairs);
    rc = run_passes(plan->inv, F + 4 * n, st, nullptr, (unsigned)nv, n);
    if (rc) return rc;
    size_t nch = (size_t)(cm + cv - 1);
    hipLaunchKernelGGL(k_overlap_add, dim3((unsigned)((ow + 255) / 256), nv), dim3(256), 0, st,
                       F + 4 * n, nch, dout, ow, (u64)n, (u64)ow);
    GF2T_CHECK_LAUNCH("gf2t_mat2n_host");
    if (!zc) GF2T_HIP(hipMemcpyAsync(hc.hbuf, dout, nv * ow * 8, hipMemcpyDeviceToHost, st));
    GF2T_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nv; i++) {
        std::memcpy(hout[i], hc.hbuf + (size_t)i * ow, ow * 8);
        if (outw > ow) std::memset(hout[i] + ow, 0, (outw - ow) * 8);
    }
    return GF2T_OK;
}

extern "C" int gf2t_mat2_host(const uint64_t *const hm[4], const uint64_t mbits[4],
                              const uint64_t *ha, uint64_t abits, const uint64_t *hb,
                              uint64_t bbits, uint64_t *hoa, uint64_t *hob, size_t outw) {
    const uint64_t *hv[2] = {ha, hb};
    const uint64_t vbits[2] = {abits, bbits};
    uint64_t *hout[2] = {hoa, hob};
    if (!hoa || !hob) {
        gf2t::set_error("gf2t_mat2_host: null pointer");
        return GF2T_EINVAL;
    }
    return gf2t_mat2n_host(hm, mbits, 1, hv, vbits, hout, outw);
}
