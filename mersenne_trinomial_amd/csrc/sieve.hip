// sieve.hip -- (a) Zech-log coarse sieve on gfx950.
//
// For each degree d: GF(2^d) = GF(2)[x]/q with q primitive, so g = x has
// order M = 2^d - 1.  An irreducible P of degree d with root beta = g^j
// divides x^r + x^s + 1 iff beta^r + beta^s = 1, i.e. j*s == Z(u) (mod M)
// with u = j*r mod M and Z the Zech logarithm Z(u) = log(g^u + 1): an
// arithmetic progression in s of modulus n = M / gcd(j, M).  One root per
// conjugacy class suffices: the class representative is the j that is
// strictly smaller than all its other rotations j*2^i mod M (a binary
// Lyndon word of length d), which also excludes every element of a
// proper subfield.  Each progression is marked into best[s] with an
// atomic min of (d << 48) | mask, i.e. smallest degree, then numerically
// least mask (reference: src/sieve_ref.cpp:1-30, src/coarse_sieve.cu:
// 237-299).
//
// Per degree, on the device:
//   k_log_build   L[g^i] = i for all i < M (LFSR chain per lane, random
//                 4-byte scatter: the HBM-bound part, 4*2^d bytes)
//   k_lyndon_t    a run of 16 candidates j per thread, all rotations
//                 compared, one output reservation per block: the class
//                 representatives as a dense list
//   k_zech        one lane per class: Zech log (one random gather) and the
//                 progression; classes whose progression reaches s <= smax
//                 are ballot-compacted into a queue (about 1.6% of them at
//                 d = 32 for smax = 68,139,920)
//   k_mark        one lane per queued class: minimal polynomial by
//                 Berlekamp-Massey on Tr(beta^k), then the marks (inline up
//                 to 1024, longer progressions to k_mark_heavy)
//   k_mark_heavy  many blocks per heavy progression
// The dense tables fit HBM for every d <= 32 (16 GiB at d = 32), so no
// bucketed low-memory mode is needed on a 288 GB part.
#include "gf2t_internal.hpp"
#include "../../include/gf2t.h"
#include <vector>
#include <algorithm>
#include <cstring>
#include <mutex>

namespace {

constexpr u64 INLINE_MARKS = 1024;
constexpr u32 HEAVY_CAP = 1u << 22;
constexpr int LOG_CHUNK = 64;

struct Heavy {
    u64 s_first, n, key;
};

// ------------------------------------------------------------ host field
u64 hclmul(u64 a, u64 b, u64 &hi) {
    u64 lo = 0;
    hi = 0;
    for (int i = 0; i < 64; i++)
        if ((b >> i) & 1) {
            lo ^= a << i;
            if (i) hi ^= a >> (64 - i);
        }
    return lo;
}
u64 hmulmod(u64 a, u64 b, u64 q, int d) {
    u64 hi, lo = hclmul(a, b, hi);
    for (int bit = 2 * d - 2; bit >= d; bit--) {
        u64 have = bit >= 64 ? (hi >> (bit - 64)) & 1 : (lo >> bit) & 1;
        if (!have) continue;
        int sh = bit - d;
        lo ^= q << sh;
        if (sh) hi ^= q >> (64 - sh);
    }
    return lo;
}
u64 hpgcd(u64 a, u64 b) {
    while (b) {
        int db = 63 - __builtin_clzll(b);
        while (a && 63 - __builtin_clzll(a) >= db) a ^= b << ((63 - __builtin_clzll(a)) - db);
        std::swap(a, b);
    }
    return a;
}
int hprimes(u64 m, u64 *ps) {
    int n = 0;
    for (u64 p = 2; p * p <= m; p += (p == 2 ? 1 : 2))
        if (m % p == 0) { ps[n++] = p; while (m % p == 0) m /= p; }
    if (m > 1) ps[n++] = m;
    return n;
}
// Rabin: x^(2^d) == x and gcd(x^(2^(d/p)) - x, q) = 1 for primes p | d
bool hirreducible(u64 q, int d) {
    u64 chain[65];
    u64 t = 2;
    chain[0] = t;
    for (int i = 1; i <= d; i++) { t = hmulmod(t, t, q, d); chain[i] = t; }
    if (chain[d] != 2) return false;
    u64 ps[16];
    int np = hprimes((u64)d, ps);
    for (int i = 0; i < np; i++)
        if (hpgcd(q, chain[d / ps[i]] ^ 2) != 1) return false;
    return true;
}
u64 hpow(u64 b, u64 e, u64 q, int d) {
    u64 r = 1;
    while (e) {
        if (e & 1) r = hmulmod(r, b, q, d);
        b = hmulmod(b, b, q, d);
        e >>= 1;
    }
    return r;
}
u64 primitive_poly(int d) {
    u64 M = ((u64)1 << d) - 1, ps[16];
    int np = hprimes(M, ps);
    for (u64 q = ((u64)1 << d) | 1; q < ((u64)2 << d); q += 2) {
        if (!hirreducible(q, d)) continue;
        bool prim = true;
        for (int i = 0; i < np && prim; i++)
            if (hpow(2, M / ps[i], q, d) == 1) prim = false;
        if (prim) return q;
    }
    return 0;
}
u64 gcd_host(u64 a, u64 b) {
    while (b) { u64 t = a % b; a = b; b = t; }
    return a;
}
int64_t mobius(u64 n) {
    int64_t mu = 1;
    for (u64 p = 2; p * p <= n; p++) {
        if (n % p) continue;
        n /= p;
        if (n % p == 0) return 0;
        mu = -mu;
    }
    if (n > 1) mu = -mu;
    return mu;
}
u64 irreducible_count(int d) {
    int64_t acc = 0;
    for (int k = 1; k <= d; k++)
        if (d % k == 0) acc += mobius((u64)k) * ((int64_t)1 << (d / k));
    return (u64)(acc / d);
}

// Per-degree host constants, computed once per process: the smallest
// primitive modulus q, the antilog tables g^k and g^(65536 k) (k < 2^16),
// and the trace mask (bit i = Tr(x^i)).  The g^(65536 k) table steps by a
// fixed multiplier, which is linear over GF(2): four 256-entry byte tables
// make each step four lookups.
struct DegHost {
    bool ready = false, trace_ok = false;
    u64 q = 0;
    u32 tmask = 0;
    std::vector<u32> lo, hi;
};
std::mutex g_deg_mu;
DegHost g_deg[33];

const DegHost &deg_host(int d) {
    std::lock_guard<std::mutex> lk(g_deg_mu);
    DegHost &H = g_deg[d];
    if (H.ready) return H;
    H.q = primitive_poly(d);
    H.lo.assign(65536, 0);
    H.hi.assign(65536, 0);
    if (H.q) {
        const u64 q = H.q;
        u64 a = 1;
        for (u64 k = 0; k < 65536; k++) {
            H.lo[k] = (u32)a;
            a <<= 1;
            if ((a >> d) & 1) a ^= q;
        }
        const u64 g16 = a;   // g^65536
        u32 T[4][256];
        for (int j = 0; j < 4; j++)
            for (u64 v = 0; v < 256; v++) {
                u64 e = v << (8 * j);
                T[j][v] = (e >> d) ? 0u : (u32)hmulmod(e, g16, q, d);
            }
        u32 b = 1;
        for (u64 k = 0; k < 65536; k++) {
            H.hi[k] = b;
            b = T[0][b & 255] ^ T[1][(b >> 8) & 255] ^ T[2][(b >> 16) & 255] ^ T[3][b >> 24];
        }
        H.trace_ok = true;
        for (int i = 0; i < d; i++) {
            u64 y = (u64)1 << i, acc = y;
            for (int k = 1; k < d; k++) { y = hmulmod(y, y, q, d); acc ^= y; }
            if (acc > 1) { H.trace_ok = false; break; }
            H.tmask |= (u32)acc << i;
        }
    }
    H.ready = true;
    return H;
}

// ---------------------------------------------------------- device side
struct DegreeParams {
    int d;
    u64 M, q, rmod, smax;
    u32 tmask;               // Tr(y) = parity(y & tmask)
    const u32 *Alo, *Ahi;    // g^k and g^(k<<16), k < 2^16
};

GF2T_DEV u32 antilog(const DegreeParams &P, u64 e) {
    if (P.d <= 16) return P.Alo[e];
    return fmul(P.Ahi[e >> 16], P.Alo[e & 0xFFFF], P.q, P.d);
}

__global__ void k_log_build(DegreeParams P, u32 *__restrict__ L) {
    u64 chunk = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 i0 = chunk * LOG_CHUNK;
    if (i0 >= P.M) return;
    u32 a = antilog(P, i0);
    const u32 top = 1u << (P.d - 1);
    // (u32)q keeps the x^d term for d < 32 (it cancels the shifted-out
    // bit) and drops it for d = 32 (the shift already discarded it)
    const u32 qlow = (u32)P.q;
    u64 iend = i0 + LOG_CHUNK < P.M ? i0 + LOG_CHUNK : P.M;
    for (u64 i = i0; i < iend; i++) {
        L[a] = (u32)i;
        a = (a << 1) ^ ((a & top) ? qlow : 0u);
    }
}

// exclusive prefix sum of v over the block (blockDim.x = T), and the total;
// wtot: T/64 words of LDS, free again after the caller's next barrier
template <int T>
__device__ unsigned block_excl_scan(unsigned v, unsigned *total, unsigned *wtot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wtot[w] = x;
    __syncthreads();
    unsigned before = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < T / 64; i++) {
        const unsigned t = wtot[i];
        if (i < w) before += t;
        tot += t;
    }
    *total = tot;
    return before + x - v;
}

// off[b] = sum of cnt[0..b) for the nb <= 8*T buckets of a block-level
// counting sort, all threads taking part (each sums a run of buckets,
// then one block scan); the caller waits at a barrier before reading off
template <int T>
__device__ void bucket_offsets(const unsigned *cnt, unsigned *off, unsigned nb, unsigned *wtot) {
    const unsigned per = (nb + T - 1) / T;
    const unsigned b0 = threadIdx.x * per;
    unsigned v = 0;
    for (unsigned i = 0; i < per; i++)
        if (b0 + i < nb) v += cnt[b0 + i];
    unsigned total;
    unsigned acc = block_excl_scan<T>(v, &total, wtot);
    for (unsigned i = 0; i < per; i++)
        if (b0 + i < nb) {
            off[b0 + i] = acc;
            acc += cnt[b0 + i];
        }
}

// ---- log table by radix partition (d >= LP_MIN_D) --------------------
// The direct build above scatters 4-byte stores over 4*2^d bytes.  For
// large d the table is built in three coalesced passes instead:
//   pass 1: each block generates E consecutive (g^i, i), sorts them in LDS
//           by the top B1 bits of g^i and appends each bucket's run to
//           that bucket's region (one cursor atomic per bucket per block);
//   pass 2: the same by the next B2 bits, inside each level-1 bucket;
//   pass 3: one block per sub-bucket of 2^LP_SB values places its entries
//           into an LDS copy of that slice of L and writes it out whole.
// Every value 1..2^d-1 occurs exactly once (g generates the group), so
// every region's size is known in advance.  Entries are u64:
// (g^i mod 2^(d-B1)) << 32 | i.
constexpr int LP_MIN_D = 26;
constexpr int LP_SB = 15;            // 2^15 values per placed slice (128 KiB)
constexpr int LP_E = 16384;          // entries per block in passes 1 and 2
constexpr int LP_T = 1024;           // threads per block in passes 1 and 2
constexpr int LP_B1 = 9;            // level-1 partition bits (fewer when d - LP_SB is smaller)

struct LpShape {
    int B1, B2;                      // partition bits, level 1 and 2
};

__global__ __launch_bounds__(LP_T) void k_lp_part1(DegreeParams P, LpShape S,
                                                   unsigned long long *__restrict__ out,
                                                   unsigned *__restrict__ cursor,
                                                   unsigned *__restrict__ bad) {
    __shared__ unsigned cnt[1 << LP_B1], off[1 << LP_B1], gbase[1 << LP_B1], wtot[LP_T / 64];
    __shared__ unsigned long long stage[LP_E];
    const unsigned nb = 1u << S.B1;
    const int shift = P.d - S.B1;
    const u64 cap = (u64)1 << shift;
    const u64 i0 = (u64)blockIdx.x * LP_E;
    for (unsigned b = threadIdx.x; b < nb; b += LP_T) cnt[b] = 0;
    __syncthreads();
    // each thread: LP_E / LP_T consecutive exponents by the LFSR step a <- a*g
    constexpr int PER = LP_E / LP_T;
    const u64 my0 = i0 + (u64)threadIdx.x * PER;
    unsigned rank[PER];
    u32 av[PER];
    const u32 top = 1u << (P.d - 1);
    const u32 qlow = (u32)P.q;
    u32 a = my0 < P.M ? antilog(P, my0) : 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        av[k] = a;
        rank[k] = 0;
        if (my0 + k < P.M) rank[k] = atomicAdd(&cnt[a >> shift], 1u);
        a = (a << 1) ^ ((a & top) ? qlow : 0u);
    }
    __syncthreads();
    bucket_offsets<LP_T>(cnt, off, nb, wtot);
    __syncthreads();
    for (unsigned b = threadIdx.x; b < nb; b += LP_T)
        gbase[b] = cnt[b] ? atomicAdd(&cursor[b], cnt[b]) : 0;
#pragma unroll
    for (int k = 0; k < PER; k++)
        if (my0 + k < P.M)
            stage[off[av[k] >> shift] + rank[k]] = ((unsigned long long)av[k] << 32) | (u32)(my0 + k);
    __syncthreads();
    const u64 nvalid = (i0 + LP_E <= P.M) ? LP_E : P.M - i0;
    for (u64 k = threadIdx.x; k < nvalid; k += LP_T) {
        unsigned long long e = stage[k];
        u32 v = (u32)(e >> 32);
        unsigned b = v >> shift;
        u64 pos = (u64)gbase[b] + (k - off[b]);
        if (pos >= cap) { atomicOr(bad, 4u); continue; }
        out[(u64)b * cap + pos] = ((unsigned long long)(v & (u32)(cap - 1)) << 32) | (u32)e;
    }
}

// level 2: inside level-1 bucket blockIdx.y, chunk blockIdx.x of LP_E
// entries, by bits [LP_SB, LP_SB + B2) of the stored value
__global__ __launch_bounds__(LP_T) void k_lp_part2(DegreeParams P, LpShape S,
                                                   const unsigned long long *__restrict__ in,
                                                   unsigned long long *__restrict__ out,
                                                   unsigned *__restrict__ cursor,
                                                   unsigned *__restrict__ bad) {
    __shared__ unsigned cnt[512], off[512], gbase[512], wtot[LP_T / 64];
    __shared__ unsigned long long stage[LP_E];
    const unsigned nb2 = 1u << S.B2;
    const u64 cap1 = (u64)1 << (P.d - S.B1);
    const u64 cap2 = (u64)1 << LP_SB;
    const unsigned b1 = blockIdx.y;
    const u64 count = b1 ? cap1 : cap1 - 1;     // value 0 never occurs
    const u64 k0 = (u64)blockIdx.x * LP_E;
    if (k0 >= count) return;
    const u64 nvalid = (k0 + LP_E <= count) ? LP_E : count - k0;
    const unsigned long long *src = in + (u64)b1 * cap1 + k0;
    for (unsigned b = threadIdx.x; b < nb2; b += LP_T) cnt[b] = 0;
    __syncthreads();
    constexpr int PER = LP_E / LP_T;
    unsigned rank[PER];
    unsigned long long ev[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        u64 idx = (u64)k * LP_T + threadIdx.x;
        ev[k] = idx < nvalid ? src[idx] : 0;
        rank[k] = idx < nvalid ? atomicAdd(&cnt[(unsigned)(ev[k] >> (32 + LP_SB)) & (nb2 - 1)], 1u) : 0;
    }
    __syncthreads();
    bucket_offsets<LP_T>(cnt, off, nb2, wtot);
    __syncthreads();
    for (unsigned b = threadIdx.x; b < nb2; b += LP_T)
        gbase[b] = cnt[b] ? atomicAdd(&cursor[(u64)b1 * nb2 + b], cnt[b]) : 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        u64 idx = (u64)k * LP_T + threadIdx.x;
        if (idx < nvalid) stage[off[(unsigned)(ev[k] >> (32 + LP_SB)) & (nb2 - 1)] + rank[k]] = ev[k];
    }
    __syncthreads();
    for (u64 k = threadIdx.x; k < nvalid; k += LP_T) {
        unsigned long long e = stage[k];
        unsigned b = (unsigned)(e >> (32 + LP_SB)) & (nb2 - 1);
        u64 pos = (u64)gbase[b] + (k - off[b]);
        if (pos >= cap2) { atomicOr(bad, 4u); continue; }
        out[((u64)b1 * nb2 + b) * cap2 + pos] = e;
    }
}

// one block per slice z of 2^LP_SB values: L[z*2^LP_SB + v] for every v
__global__ __launch_bounds__(1024) void k_lp_place(const unsigned long long *__restrict__ in,
                                                   u32 *__restrict__ L) {
    extern __shared__ u32 tile[];
    const u64 cap = (u64)1 << LP_SB;
    const unsigned z = blockIdx.x;
    const u64 count = z ? cap : cap - 1;
    const unsigned long long *src = in + (u64)z * cap;
    if (z == 0 && threadIdx.x == 0) tile[0] = 0;
    for (u64 k = threadIdx.x; k < count; k += blockDim.x) {
        unsigned long long e = src[k];
        tile[(u32)(e >> 32) & (u32)(cap - 1)] = (u32)e;
    }
    __syncthreads();
    u32 *dst = L + (u64)z * cap;
    for (u64 k = threadIdx.x; k < cap; k += blockDim.x) dst[k] = tile[k];
}

__global__ void k_fill_u32(unsigned *p, u64 n) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

LpShape lp_shape(int d) {
    LpShape S;
    S.B1 = std::min(LP_B1, d - LP_SB);
    S.B2 = d - LP_SB - S.B1;
    return S;
}

// A Lyndon word of length >= 2 starts with 0 and ends with 1, so only the
// odd j below 2^(d-1) are candidates: candidate c tests j = 2c + 1,
// c < 2^(d-2).  Each thread tests a run of LY_PER consecutive candidates
// against all d-1 rotations (no early exit: about one candidate in d/4 is
// a Lyndon word, so nearly every wave runs all rotations anyway; the
// LY_PER independent chains keep the SIMD busy), the block sums its
// counts, and one atomic per block reserves the block's output run.
// Survivors stay ascending within each block's run.
constexpr int LY_T = 256, LY_PER = 16;

template <class W>
__global__ __launch_bounds__(LY_T) void k_lyndon_t(int d, W M, W *__restrict__ surv,
                                                   unsigned long long *__restrict__ nsurv,
                                                   u64 cap) {
    __shared__ unsigned wtot[LY_T / 64];
    __shared__ unsigned long long bbase;
    const u64 ncand = (u64)1 << (d - 2);
    const u64 chunk = (u64)LY_T * LY_PER;
    for (u64 c0 = (u64)blockIdx.x * chunk; c0 < ncand; c0 += (u64)gridDim.x * chunk) {
        const u64 cb = c0 + (u64)threadIdx.x * LY_PER;
        W jv[LY_PER];
        bool ok[LY_PER];
#pragma unroll
        for (int i = 0; i < LY_PER; i++) {
            jv[i] = (W)(2 * (cb + i) + 1);
            ok[i] = cb + i < ncand;
        }
        for (int k = 1; k < d; k++) {
#pragma unroll
            for (int i = 0; i < LY_PER; i++)
                ok[i] = ok[i] && ((((jv[i] << k) | (jv[i] >> (d - k))) & M) > jv[i]);
        }
        unsigned keep = 0;
#pragma unroll
        for (int i = 0; i < LY_PER; i++) keep |= (unsigned)ok[i] << i;
        unsigned total;
        const unsigned pre = block_excl_scan<LY_T>((unsigned)__popc(keep), &total, wtot);
        if (threadIdx.x == 0) bbase = total ? atomicAdd(nsurv, (unsigned long long)total) : 0;
        __syncthreads();
        u64 pos = bbase + pre;
#pragma unroll
        for (int i = 0; i < LY_PER; i++)
            if ((keep >> i) & 1) {
                if (pos < cap) surv[pos] = jv[i];
                pos++;
            }
        __syncthreads();   // bbase and wtot are rewritten by the next chunk
    }
}

unsigned lyndon_blocks(int d) {
    const u64 chunks = ((((u64)1 << (d - 2)) + (u64)LY_T * LY_PER - 1) / ((u64)LY_T * LY_PER));
    return (unsigned)std::min<u64>(chunks, 8192);
}

struct MarkRec {
    u32 j, n, sf, pad;
};

// Phase 1, one lane per class: Zech logarithm (one random gather), the
// progression s == s0 (mod n), and -- only for the few classes whose
// progression reaches s <= smax -- a ballot-compacted queue entry, so the
// expensive minimal polynomial below runs on dense waves.
constexpr int ZC_T = 256, ZC_PER = 4;   // classes per thread per round

__global__ __launch_bounds__(ZC_T) void k_zech(DegreeParams P, const u32 *__restrict__ surv,
                                               u64 nsurv, const u32 *__restrict__ L,
                                               MarkRec *__restrict__ queue,
                                               unsigned long long *__restrict__ nqueue, u64 qcap) {
    __shared__ unsigned wtot[ZC_T / 64];
    __shared__ unsigned long long bbase;
    const u32 M = (u32)P.M;
    const u64 chunk = (u64)ZC_T * ZC_PER;
    // a round: ZC_PER classes per thread, their table lookups all issued
    // before any of the arithmetic, then one queue reservation per block
    for (u64 c0 = (u64)blockIdx.x * chunk; c0 < nsurv; c0 += (u64)gridDim.x * chunk) {
        u32 j[ZC_PER], a[ZC_PER], v[ZC_PER];
        bool live[ZC_PER];
#pragma unroll
        for (int q = 0; q < ZC_PER; q++) {
            const u64 t = c0 + (u64)q * ZC_T + threadIdx.x;
            live[q] = t < nsurv;
            j[q] = live[q] ? surv[t] : 1u;
        }
#pragma unroll
        for (int q = 0; q < ZC_PER; q++) {
            const u64 u = mod_mersenne((u64)j[q] * P.rmod, P.d, P.M);
            live[q] = live[q] && u;                    // u = 0: beta^r = 1, never a root
            a[q] = antilog(P, u ? u : 1) ^ 1u;
        }
#pragma unroll
        for (int q = 0; q < ZC_PER; q++) v[q] = live[q] ? L[a[q]] : 0u;   // Zech logarithms of u
        unsigned keep = 0;
        MarkRec rec[ZC_PER];
#pragma unroll
        for (int q = 0; q < ZC_PER; q++) {
            rec[q] = MarkRec{0, 0, 0, 0};
            if (!live[q]) continue;
            const u32 gg = gcd32(j[q], M);
            if (v[q] % gg) continue;
            const u32 n = M / gg;
            const u32 iv = inv32(j[q] / gg, n);
            const u64 prod = (u64)iv * (v[q] / gg);
            const u32 s0 = (u32)(gg == 1 ? mod_mersenne(prod, P.d, P.M) : prod % n);
            const u32 sf = s0 ? s0 : n;
            if (sf <= P.smax) {
                keep |= 1u << q;
                rec[q] = MarkRec{j[q], n, sf, 0};
            }
        }
        unsigned total;
        const unsigned pre = block_excl_scan<ZC_T>((unsigned)__popc(keep), &total, wtot);
        if (threadIdx.x == 0) bbase = total ? atomicAdd(nqueue, (unsigned long long)total) : 0;
        __syncthreads();
        u64 pos = bbase + pre;
#pragma unroll
        for (int q = 0; q < ZC_PER; q++)
            if ((keep >> q) & 1) {
                if (pos < qcap) queue[pos] = rec[q];
                pos++;
            }
        __syncthreads();   // bbase and wtot are rewritten by the next round
    }
}

// Phase 2, one lane per queued class: minimal polynomial by
// Berlekamp-Massey on the trace sequence, then the marks (inline up to
// INLINE_MARKS, longer progressions to the heavy queue).
__global__ void k_mark(DegreeParams P, const MarkRec *__restrict__ queue, u64 nq,
                       unsigned long long *__restrict__ best, Heavy *__restrict__ heavy,
                       unsigned *__restrict__ nheavy, unsigned *__restrict__ bad) {
    u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const MarkRec rec = queue[t];
    const int d = P.d;
    u64 mask = minpoly_bm(antilog(P, rec.j), P.q, d, P.tmask);
    if (mask == 0 || !((mask >> d) & 1) || !(mask & 1)) { atomicAdd(bad, 1u); return; }
    unsigned long long key = ((unsigned long long)d << 48) | mask;
    u64 n = rec.n, sf = rec.sf;
    u64 count = (P.smax - sf) / n + 1;
    if (count <= INLINE_MARKS) {
        for (u64 s = sf; s <= P.smax; s += n)
            if (best[s] > key) atomicMin(&best[s], key);
    } else {
        unsigned idx = atomicAdd(nheavy, 1u);
        if (idx < HEAVY_CAP) heavy[idx] = Heavy{sf, n, key};
        else atomicAdd(bad, 2u);
    }
}

__global__ void k_mark_heavy(const Heavy *__restrict__ heavy, unsigned nheavy, u64 smax,
                             unsigned long long *__restrict__ best) {
    for (unsigned rec = blockIdx.y; rec < nheavy; rec += gridDim.y) {
        const Heavy h = heavy[rec];
        u64 count = (smax - h.s_first) / h.n + 1;
        for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < count;
             k += (u64)gridDim.x * blockDim.x) {
            u64 s = h.s_first + k * h.n;
            if (best[s] > h.key) atomicMin(&best[s], (unsigned long long)h.key);
        }
    }
}

__global__ void k_zero_counters(unsigned long long *a, unsigned *b) {
    if (threadIdx.x < 4) a[threadIdx.x] = 0;
    if (threadIdx.x < 4) b[threadIdx.x] = 0;
}


struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t b) {
        if (bytes >= b) return GF2T_OK;
        if (p) GF2T_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
        GF2T_HIP(hipMalloc(&p, b));
        bytes = b;
        return GF2T_OK;
    }
};

// ===================================================== degrees 33..36
// Elements of GF(2^d) are u64.  The log table holds d-bit exponents, so
// it is stored as the low 32 bits (L_lo, u32) plus the top four bits
// packed eight to a word (L_hi).  It is built by the same radix
// partition as above, except that the level-2 split runs one level-1
// bucket at a time, to a bucket-sized scratch region and back (two
// counting-sort passes), so the whole build needs 8 bytes per element
// plus the table (about 205 GB at d = 34, inside 288 GB).  Entries are
// (g^i mod 2^(d-8)) << 36 | i.
// Above d = 34 the table is built and used in value slices: slice k holds
// the logs of the field elements whose top byte lies in [k nb, (k+1) nb),
// nb = 256 / slices.  Each slice's partition pass generates every power
// but keeps only its own buckets, and the Zech pass of the slice resolves
// only the classes whose lookup value falls in it; the marking queue
// collects all slices' classes.  So the memory of d = 35 (two slices) and
// d = 36 (four) is that of d = 34, and the extra cost is one generation
// pass over the powers per slice.
constexpr int WIDE_MAX_D = 36;
constexpr int W_DENSE_D = 34;       // largest degree with the whole table at once
constexpr int W_SB = 15;            // values per placed slice: 2^15
constexpr int W_E = 8192;           // entries per block in passes 1 and 2
constexpr int W_T = 256;
constexpr int W_IB = 36;            // bits of i in an entry
constexpr int W_B1 = 8;             // level-1 bucket bits
constexpr int W_AB = 18;            // antilog table split: g^e = A1[e >> 18] * A0[e & (2^18 - 1)]

struct WParams {
    int d;
    u64 M, q, rmod, smax, tmask;
    const u64 *A0, *A1;
};

GF2T_DEV u64 antilog_w(const WParams &P, u64 e) {
    return fmul64(P.A1[e >> W_AB], P.A0[e & ((1u << W_AB) - 1)], P.q, P.d);
}

// block-wide exclusive scan of cnt[0..nb) (nb <= W_T * 16) into off;
// part: W_T / 64 words of LDS
__device__ void block_scan(const unsigned *cnt, unsigned *off, unsigned nb, unsigned *part) {
    bucket_offsets<W_T>(cnt, off, nb, part);
    __syncthreads();
}

// Pass 1 of the table build for the level-1 buckets [b0, b0 + nb): every
// power g^i is generated, and those whose top byte is in range are
// appended to bucket (top byte - b0) of out (2^(d-8) entries each).
__global__ __launch_bounds__(W_T) void k_w_part1(WParams P, unsigned long long *__restrict__ out,
                                                 unsigned *__restrict__ cursor,
                                                 unsigned *__restrict__ bad, unsigned b0,
                                                 unsigned nb) {
    __shared__ unsigned cnt[256], off[256], gbase[256], part[W_T];
    __shared__ unsigned long long stage[W_E];
    __shared__ unsigned char sbkt[W_E];   // bucket of each staged entry
    const int shift = P.d - W_B1;
    const u64 cap = (u64)1 << shift;
    const u64 i0 = (u64)blockIdx.x * W_E;
    cnt[threadIdx.x] = 0;
    __syncthreads();
    constexpr int PER = W_E / W_T;
    const u64 my0 = i0 + (u64)threadIdx.x * PER;
    unsigned rank[PER];
    u64 a0 = my0 < P.M ? antilog_w(P, my0) : 0, a = a0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        rank[k] = 0;
        const unsigned b = (unsigned)(a >> shift);
        if (my0 + k < P.M && b - b0 < nb) rank[k] = atomicAdd(&cnt[b], 1u);
        a <<= 1;
        if (a >> P.d) a ^= P.q;
    }
    __syncthreads();
    block_scan(cnt, off, 256, part);
    gbase[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]) : 0;
    a = a0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const unsigned b = (unsigned)(a >> shift);
        if (my0 + k < P.M && b - b0 < nb) {
            const unsigned at = off[b] + rank[k];
            stage[at] = ((unsigned long long)(a & (cap - 1)) << W_IB) | (my0 + k);
            sbkt[at] = (unsigned char)b;
        }
        a <<= 1;
        if (a >> P.d) a ^= P.q;
    }
    __syncthreads();
    const u64 nvalid = (u64)off[255] + cnt[255];   // entries staged in range
    for (u64 k = threadIdx.x; k < nvalid; k += W_T) {
        const unsigned b = sbkt[k];
        u64 pos = (u64)gbase[b] + (k - off[b]);
        if (pos >= cap) { atomicOr(bad, 4u); continue; }
        out[(u64)(b - b0) * cap + pos] = stage[k];
    }
}

// One counting-sort pass of the level-2 split, which runs as two such
// passes per level-1 bucket (about half the d - 23 level-2 bits each, so
// every block writes runs of 128-256 entries): source bucket blockIdx.y
// (src_cap entries apart; the first holds first_count) is partitioned by
// the nbits-bit digit at bit `shift` of the stored value into
// destination buckets src * 2^nbits + digit of dst_cap entries each.
constexpr int WS_MAXB = 128;  // digits per pass (2^7: degree 36 splits 13 bits as 7 + 6)
constexpr int WS_T = 512;     // threads per block
__global__ __launch_bounds__(WS_T) void k_w_split(const unsigned long long *__restrict__ in,
                                                 u64 src_cap, u64 first_count, int shift,
                                                 int nbits, unsigned long long *__restrict__ out,
                                                 u64 dst_cap, unsigned *__restrict__ cursor,
                                                 unsigned *__restrict__ bad) {
    __shared__ unsigned cnt[WS_MAXB], off[WS_MAXB], gbase[WS_MAXB], part[WS_T / 64];
    __shared__ unsigned long long stage[W_E];
    const unsigned nd = 1u << nbits;
    const unsigned src = blockIdx.y;
    const u64 count = src ? src_cap : first_count;
    const u64 k0 = (u64)blockIdx.x * W_E;
    if (k0 >= count) return;
    const u64 nvalid = (k0 + W_E <= count) ? W_E : count - k0;
    const unsigned long long *sp = in + (u64)src * src_cap + k0;
    const int dsh = W_IB + shift;
    for (unsigned b = threadIdx.x; b < nd; b += WS_T) cnt[b] = 0;
    __syncthreads();
    constexpr int PER = W_E / WS_T;
    unsigned rank[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const u64 idx = (u64)k * WS_T + threadIdx.x;
        rank[k] = idx < nvalid ? atomicAdd(&cnt[(unsigned)(sp[idx] >> dsh) & (nd - 1)], 1u) : 0;
    }
    __syncthreads();
    bucket_offsets<WS_T>(cnt, off, nd, part);
    __syncthreads();
    for (unsigned b = threadIdx.x; b < nd; b += WS_T)
        gbase[b] = cnt[b] ? atomicAdd(&cursor[src * nd + b], cnt[b]) : 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const u64 idx = (u64)k * WS_T + threadIdx.x;
        if (idx < nvalid) {
            const unsigned long long e = sp[idx];
            stage[off[(unsigned)(e >> dsh) & (nd - 1)] + rank[k]] = e;
        }
    }
    __syncthreads();
    for (u64 k = threadIdx.x; k < nvalid; k += WS_T) {
        const unsigned long long e = stage[k];
        const unsigned b = (unsigned)(e >> dsh) & (nd - 1);
        const u64 pos = (u64)gbase[b] + (k - off[b]);
        if (pos >= dst_cap) { atomicOr(bad, 4u); continue; }
        out[((u64)src * nd + b) * dst_cap + pos] = e;
    }
}

// one block per run z of 2^W_SB values of the slice: L_lo and the
// packed top bits.  zero_first: the slice starts at value 0, which has no
// logarithm (its run holds 2^W_SB - 1 entries; L[0] is left 0)
__global__ __launch_bounds__(1024) void k_w_place(const unsigned long long *__restrict__ in,
                                                  u32 *__restrict__ Llo, u32 *__restrict__ Lhi,
                                                  int zero_first) {
    extern __shared__ u32 wtile[];
    u32 *lo = wtile, *hi = wtile + (1u << W_SB);          // hi: 4 bits per value
    const u64 cap = (u64)1 << W_SB;
    const u64 z = blockIdx.x;
    const bool has_zero = zero_first && z == 0;
    const u64 count = has_zero ? cap - 1 : cap;
    const unsigned long long *src = in + z * cap;
    for (u64 k = threadIdx.x; k < cap / 8; k += blockDim.x) hi[k] = 0;
    if (has_zero && threadIdx.x == 0) lo[0] = 0;
    __syncthreads();
    for (u64 k = threadIdx.x; k < count; k += blockDim.x) {
        unsigned long long e = src[k];
        u32 v = (u32)(e >> W_IB) & (u32)(cap - 1);
        u64 i = e & (((u64)1 << W_IB) - 1);
        lo[v] = (u32)i;
        atomicOr(&hi[v >> 3], (u32)((i >> 32) & 15) << (4 * (v & 7)));
    }
    __syncthreads();
    for (u64 k = threadIdx.x; k < cap; k += blockDim.x) Llo[z * cap + k] = lo[k];
    for (u64 k = threadIdx.x; k < cap / 8; k += blockDim.x) Lhi[z * (cap / 8) + k] = hi[k];
}

struct MarkRecW {
    u64 j, n, sf;
};

// as k_zech: ZC_PER classes per thread, lookups first, one reservation
// per block; only the classes whose lookup value y = beta^r + 1 lies in
// the table slice [ybase, ybase + yspan) are resolved
__global__ __launch_bounds__(ZC_T) void k_zech_w(WParams P, const u64 *__restrict__ surv,
                                                 u64 nsurv, const u32 *__restrict__ Llo,
                                                 const u32 *__restrict__ Lhi,
                                                 MarkRecW *__restrict__ queue,
                                                 unsigned long long *__restrict__ nqueue,
                                                 u64 qcap, u64 ybase, u64 yspan) {
    __shared__ unsigned wtot[ZC_T / 64];
    __shared__ unsigned long long bbase;
    const u64 chunk = (u64)ZC_T * ZC_PER;
    for (u64 c0 = (u64)blockIdx.x * chunk; c0 < nsurv; c0 += (u64)gridDim.x * chunk) {
        u64 j[ZC_PER], y[ZC_PER], v[ZC_PER];
        bool live[ZC_PER];
#pragma unroll
        for (int q = 0; q < ZC_PER; q++) {
            const u64 t = c0 + (u64)q * ZC_T + threadIdx.x;
            live[q] = t < nsurv;
            j[q] = live[q] ? surv[t] : 1;
        }
#pragma unroll
        for (int q = 0; q < ZC_PER; q++) {
            const u64 u = mulmod64(j[q], P.rmod, P.M);
            live[q] = live[q] && u;                    // u = 0: beta^r = 1, never a root
            y[q] = (antilog_w(P, u ? u : 1) ^ 1u) - ybase;
            live[q] = live[q] && y[q] < yspan;         // resolved by this slice's table
        }
#pragma unroll
        for (int q = 0; q < ZC_PER; q++)
            v[q] = live[q] ? ((u64)Llo[y[q]] |
                              ((u64)((Lhi[y[q] >> 3] >> (4 * (y[q] & 7))) & 15) << 32))
                           : 0;
        unsigned keep = 0;
        MarkRecW rec[ZC_PER];
#pragma unroll
        for (int q = 0; q < ZC_PER; q++) {
            rec[q] = MarkRecW{0, 0, 0};
            if (!live[q]) continue;
            const u64 gg = gcd64(j[q], P.M);
            if (v[q] % gg) continue;
            const u64 n = P.M / gg;
            const u64 iv = inv64((j[q] / gg) % n, n);
            const u64 s0 = mulmod64(iv, v[q] / gg, n);
            const u64 sf = s0 ? s0 : n;
            if (sf <= P.smax) {
                keep |= 1u << q;
                rec[q] = MarkRecW{j[q], n, sf};
            }
        }
        unsigned total;
        const unsigned pre = block_excl_scan<ZC_T>((unsigned)__popc(keep), &total, wtot);
        if (threadIdx.x == 0) bbase = total ? atomicAdd(nqueue, (unsigned long long)total) : 0;
        __syncthreads();
        u64 pos = bbase + pre;
#pragma unroll
        for (int q = 0; q < ZC_PER; q++)
            if ((keep >> q) & 1) {
                if (pos < qcap) queue[pos] = rec[q];
                pos++;
            }
        __syncthreads();
    }
}

__global__ void k_mark_w(WParams P, const MarkRecW *__restrict__ queue, u64 nq,
                         unsigned long long *__restrict__ best, Heavy *__restrict__ heavy,
                         unsigned *__restrict__ nheavy, unsigned *__restrict__ bad) {
    u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const MarkRecW rec = queue[t];
    const int d = P.d;
    u64 mask = minpoly_bm64(antilog_w(P, rec.j), P.q, d, P.tmask);
    if (mask == 0 || !((mask >> d) & 1) || !(mask & 1)) { atomicAdd(bad, 1u); return; }
    unsigned long long key = ((unsigned long long)d << 48) | mask;
    u64 count = (P.smax - rec.sf) / rec.n + 1;
    if (count <= INLINE_MARKS) {
        for (u64 s = rec.sf; s <= P.smax; s += rec.n)
            if (best[s] > key) atomicMin(&best[s], key);
    } else {
        unsigned idx = atomicAdd(nheavy, 1u);
        if (idx < HEAVY_CAP) heavy[idx] = Heavy{rec.sf, rec.n, key};
        else atomicAdd(bad, 2u);
    }
}

struct WideHost {
    bool ready = false, trace_ok = false;
    u64 q = 0, tmask = 0;
    std::vector<u64> a0, a1;
};
std::mutex g_wide_mu;
WideHost g_wide[WIDE_MAX_D + 1];

const WideHost &wide_host(int d) {
    std::lock_guard<std::mutex> lk(g_wide_mu);
    WideHost &H = g_wide[d];
    if (H.ready) return H;
    H.q = primitive_poly(d);
    const size_t na = (size_t)1 << W_AB;
    H.a0.assign(na, 0);
    H.a1.assign(na, 0);
    if (H.q) {
        const u64 q = H.q;
        u64 a = 1;
        for (size_t k = 0; k < na; k++) {
            H.a0[k] = a;
            a <<= 1;
            if ((a >> d) & 1) a ^= q;
        }
        const u64 gstep = a;   // g^(2^18); multiplying by it is linear: byte tables
        u64 T[5][256];
        for (int j = 0; j < 5; j++)
            for (u64 v = 0; v < 256; v++) {
                u64 e = v << (8 * j);
                T[j][v] = (e >> d) ? 0 : hmulmod(e, gstep, q, d);
            }
        u64 b = 1;
        for (size_t k = 0; k < na; k++) {
            H.a1[k] = b;
            b = T[0][b & 255] ^ T[1][(b >> 8) & 255] ^ T[2][(b >> 16) & 255] ^
                T[3][(b >> 24) & 255] ^ T[4][(b >> 32) & 255];
        }
        H.trace_ok = true;
        for (int i = 0; i < d; i++) {
            u64 y = (u64)1 << i, acc = y;
            for (int k = 1; k < d; k++) { y = hmulmod(y, y, q, d); acc ^= y; }
            if (acc > 1) { H.trace_ok = false; break; }
            H.tmask |= acc << i;
        }
    }
    H.ready = true;
    return H;
}

struct SieveWork {
    DevBuf bL, bSurv, bQueue, bHeavy, bTab, bCnt, bP1, bP2, bCur;
    DevBuf wP1, wTmp, wLlo, wLhi, wSurv, wQueue, wTab, wCur;   // degrees 33..34
};
std::mutex g_sw_mu;
SieveWork *g_sw = nullptr;   // never destroyed at exit: the runtime may be gone by then

}  // namespace

extern "C" int gf2t_sieve_release(void) {
    std::lock_guard<std::mutex> lk(g_sw_mu);
    if (g_sw) {
        for (DevBuf *b : {&g_sw->bL, &g_sw->bSurv, &g_sw->bQueue, &g_sw->bHeavy, &g_sw->bTab,
                          &g_sw->bCnt, &g_sw->bP1, &g_sw->bP2, &g_sw->bCur, &g_sw->wP1,
                          &g_sw->wTmp, &g_sw->wLlo, &g_sw->wLhi, &g_sw->wSurv, &g_sw->wQueue,
                          &g_sw->wTab, &g_sw->wCur})
            b->release();
    }
    return GF2T_OK;
}

extern "C" uint64_t gf2t_primitive_poly(int d) {
    if (d < 2 || d > 40) return 0;
    return primitive_poly(d);
}

namespace {
// per-degree phase timers, destroyed on every return path
struct SieveEvents {
    hipEvent_t e[4] = {};
    ~SieveEvents() {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// Device memory of the degrees <= 32: the log table, the two partition
// buffers and the class / marking arrays of the largest degree.
struct NarrowMem {
    u32 *L;
    unsigned long long *P1, *P2;
    u32 *surv;
    MarkRec *queue;
};

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The degree-<=32 arrays are carved out of the wide tables whenever those
// are allocated (they are larger: 12 bytes per element of GF(2^33) against
// 12 per element of GF(2^32) plus the class arrays), so a sieve to degree
// 34 never holds both sets and never frees and reallocates between them.
int narrow_mem(int dn, NarrowMem &m) {
    SieveWork &W = *g_sw;
    const size_t N = (size_t)1 << dn;
    const u64 ncls = irreducible_count(dn);
    const size_t lbytes = align256(N * sizeof(u32)), sbytes = align256((ncls + 64) * sizeof(u32));
    const size_t qbytes = (ncls + 64) * sizeof(MarkRec);
    const bool part = dn >= LP_MIN_D;
    const size_t pbytes = part ? N * 8 : 0;
    if (W.wP1.bytes >= 2 * pbytes && W.wLlo.bytes >= lbytes + sbytes + qbytes) {
        char *l = (char *)W.wLlo.p;
        m.L = (u32 *)l;
        m.surv = (u32 *)(l + lbytes);
        m.queue = (MarkRec *)(l + lbytes + sbytes);
        m.P1 = (unsigned long long *)W.wP1.p;
        m.P2 = (unsigned long long *)((char *)W.wP1.p + pbytes);
        return GF2T_OK;
    }
    int rc;
    if ((rc = W.bL.ensure(lbytes))) return rc;
    if ((rc = W.bSurv.ensure(sbytes))) return rc;
    if ((rc = W.bQueue.ensure(qbytes))) return rc;
    if (part) {
        if ((rc = W.bP1.ensure(pbytes))) return rc;
        if (lp_shape(dn).B2 > 0 && (rc = W.bP2.ensure(pbytes))) return rc;
    }
    m.L = (u32 *)W.bL.p;
    m.surv = (u32 *)W.bSurv.p;
    m.queue = (MarkRec *)W.bQueue.p;
    m.P1 = (unsigned long long *)W.bP1.p;
    m.P2 = (unsigned long long *)W.bP2.p;
    return GF2T_OK;
}

// degrees dmin..dmax <= 32 (u32 field elements); g_sw_mu held by the caller
int sieve_narrow(uint64_t r, int dmin, int dmax, uint64_t smax, uint64_t *dbest,
                 gf2t_sieve_stats *stats) {
    DevBuf &bHeavy = g_sw->bHeavy, &bTab = g_sw->bTab, &bCnt = g_sw->bCnt, &bCur = g_sw->bCur;
    int rc;
    NarrowMem mem;
    if ((rc = narrow_mem(dmax, mem))) return rc;
    if (dmax >= LP_MIN_D) {
        LpShape S = lp_shape(dmax);
        if ((rc = bCur.ensure(((size_t)1 << (S.B1 + S.B2)) * 4 + (1 << LP_B1) * 4))) return rc;
        static std::once_flag lds_once;
        static hipError_t lds_err = hipSuccess;
        std::call_once(lds_once, [] {
            lds_err = hipFuncSetAttribute((const void *)k_lp_place,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)(((size_t)1 << LP_SB) * sizeof(u32)));
        });
        GF2T_HIP(lds_err);
    }
    if ((rc = bTab.ensure(2 * 65536 * sizeof(u32)))) return rc;
    if ((rc = bHeavy.ensure((size_t)HEAVY_CAP * sizeof(Heavy)))) return rc;
    if ((rc = bCnt.ensure(64))) return rc;
    unsigned long long *cnt64 = (unsigned long long *)bCnt.p;   // [0] nsurv [1] nqueue
    unsigned *cnt32 = (unsigned *)((char *)bCnt.p + 32);         // [0] nheavy [1] bad
    u32 *Alo = (u32 *)bTab.p, *Ahi = Alo + 65536;
    SieveEvents evs;
    hipEvent_t(&ev)[4] = evs.e;
    for (auto &e : ev) GF2T_HIP(hipEventCreate(&e));

    int status = GF2T_OK;
    for (int d = dmin; d <= dmax && status == GF2T_OK; d++) {
        const u64 M = ((u64)1 << d) - 1;
        const DegHost &H = deg_host(d);
        const u64 q = H.q;
        const u64 rmod = r % M;
        if (!q || gcd_host(rmod ? rmod : M, M) != 1) {
            gf2t::set_error("gf2t_sieve: d=%d: gcd(r, 2^d-1) != 1, unsupported r", d);
            status = GF2T_EINVAL;
            break;
        }
        if (!H.trace_ok) {
            gf2t::set_error("gf2t_sieve: d=%d: trace not in GF(2)", d);
            status = GF2T_EINTERNAL;
            break;
        }
        GF2T_HIP(hipMemcpy(Alo, H.lo.data(), 65536 * 4, hipMemcpyHostToDevice));
        GF2T_HIP(hipMemcpy(Ahi, H.hi.data(), 65536 * 4, hipMemcpyHostToDevice));
        const u32 tmask = H.tmask;
        DegreeParams P{d, M, q, rmod, smax, tmask, Alo, Ahi};

        const u64 ncls = irreducible_count(d);
        u32 *surv = mem.surv;
        u32 *L = mem.L;
        Heavy *heavy = (Heavy *)bHeavy.p;

        GF2T_HIP(hipEventRecord(ev[0], 0));
        if (d >= LP_MIN_D) {
            const LpShape S = lp_shape(d);
            const unsigned nb1 = 1u << S.B1, nb2 = 1u << S.B2;
            unsigned *cur1 = (unsigned *)bCur.p, *cur2 = cur1 + (1 << LP_B1);
            const u64 ncur = (1 << LP_B1) + (u64)nb1 * nb2;
            hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((ncur + 255) / 256)), dim3(256), 0, 0,
                               cur1, ncur);
            hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, 0, cnt64, cnt32);
            unsigned long long *p1 = mem.P1;
            unsigned long long *p2 = S.B2 ? mem.P2 : p1;
            hipLaunchKernelGGL(k_lp_part1, dim3((unsigned)((M + LP_E - 1) / LP_E)), dim3(LP_T), 0,
                               0, P, S, p1, cur1, cnt32 + 1);
            if (S.B2)
                hipLaunchKernelGGL(k_lp_part2,
                                   dim3((unsigned)((((u64)1 << (d - S.B1)) + LP_E - 1) / LP_E), nb1),
                                   dim3(LP_T), 0, 0, P, S, p1, p2, cur2, cnt32 + 1);
            hipLaunchKernelGGL(k_lp_place, dim3(nb1 * nb2), dim3(1024),
                               ((size_t)1 << LP_SB) * sizeof(u32), 0, p2, L);
            GF2T_CHECK_LAUNCH("log table partition");
            unsigned chk[2];
            GF2T_HIP(hipMemcpy(chk, cnt32, 8, hipMemcpyDeviceToHost));
            if (chk[1]) {
                gf2t::set_error("gf2t_sieve: d=%d: log table partition overflow", d);
                status = GF2T_EINTERNAL;
                break;
            }
        } else {
            u64 nchunks = (M + LOG_CHUNK - 1) / LOG_CHUNK;
            hipLaunchKernelGGL(k_log_build, dim3((unsigned)((nchunks + 255) / 256)), dim3(256), 0,
                               0, P, L);
            GF2T_CHECK_LAUNCH("k_log_build");
        }
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, 0, cnt64, cnt32);
        GF2T_HIP(hipEventRecord(ev[1], 0));
        hipLaunchKernelGGL(k_lyndon_t<u32>, dim3(lyndon_blocks(d)), dim3(LY_T), 0, 0, d, (u32)M,
                           surv, cnt64, (u64)(ncls + 64));
        GF2T_CHECK_LAUNCH("k_lyndon");
        unsigned long long hs[2];
        GF2T_HIP(hipMemcpy(hs, cnt64, 16, hipMemcpyDeviceToHost));
        u64 nsurv = hs[0];
        if (nsurv != ncls) {
            gf2t::set_error("gf2t_sieve: d=%d: %llu classes, expected %llu", d,
                            (unsigned long long)nsurv, (unsigned long long)ncls);
            status = GF2T_EINTERNAL;
            break;
        }
        MarkRec *queue = mem.queue;
        if (nsurv) {
            const u64 zb = std::min<u64>((nsurv + (u64)ZC_T * ZC_PER - 1) / ((u64)ZC_T * ZC_PER),
                                         1u << 16);
            hipLaunchKernelGGL(k_zech, dim3((unsigned)zb), dim3(ZC_T), 0, 0, P, surv, nsurv, L,
                               queue, cnt64 + 1, (u64)(ncls + 64));
            GF2T_CHECK_LAUNCH("k_zech");
            GF2T_HIP(hipMemcpy(hs, cnt64, 16, hipMemcpyDeviceToHost));
            u64 nq = std::min<u64>(hs[1], ncls + 64);
            if (nq) {
                hipLaunchKernelGGL(k_mark, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, 0, P,
                                   queue, nq, (unsigned long long *)dbest, heavy, cnt32,
                                   cnt32 + 1);
                GF2T_CHECK_LAUNCH("k_mark");
            }
        }
        GF2T_HIP(hipEventRecord(ev[2], 0));
        unsigned h32[2];
        GF2T_HIP(hipMemcpy(h32, cnt32, 8, hipMemcpyDeviceToHost));
        if (h32[1]) {
            gf2t::set_error("gf2t_sieve: d=%d: %s", d,
                            (h32[1] & 1) ? "minimal polynomial not over GF(2)"
                                         : "heavy queue overflow");
            status = GF2T_EINTERNAL;
            break;
        }
        unsigned nheavy = h32[0];
        if (nheavy) {
            unsigned gy = nheavy < 65535 ? nheavy : 65535;
            hipLaunchKernelGGL(k_mark_heavy, dim3(64, gy), dim3(256), 0, 0, heavy, nheavy, smax,
                               (unsigned long long *)dbest);
            GF2T_CHECK_LAUNCH("k_mark_heavy");
        }
        GF2T_HIP(hipEventRecord(ev[3], 0));
        GF2T_HIP(hipEventSynchronize(ev[3]));
        if (stats) {
            gf2t_sieve_stats &S = stats[d];
            S.d = d;
            S.q = q;
            S.classes = nsurv;
            S.marking = hs[1];
            S.heavy = nheavy;
            GF2T_HIP(hipEventElapsedTime(&S.ms_table, ev[0], ev[1]));
            GF2T_HIP(hipEventElapsedTime(&S.ms_classes, ev[1], ev[2]));
            GF2T_HIP(hipEventElapsedTime(&S.ms_heavy, ev[2], ev[3]));
        }
    }
    return status;
}

// slices of the degree-d log table: the whole table up to W_DENSE_D, then
// the fewest slices that keep a slice at 2^W_DENSE_D values; at least
// g_min_slices (a test knob, gf2t_tune "sieve_slices")
int g_min_slices = 1;
int g_wide_from = 33;   // first degree of the 64-bit path (gf2t_tune "sieve_wide_from")
int slices_for(int d) {
    int n = d > W_DENSE_D ? 1 << (d - W_DENSE_D) : 1;
    while (n < g_min_slices && n < 256) n <<= 1;
    return n;
}

// the tables of the 64-bit path up to degree dmax (one slice's worth)
int wide_mem(int dmax) {
    SieveWork &W = *g_sw;
    int rc;
    const u64 N = (u64)1 << std::min(dmax, W_DENSE_D);          // values per slice, at most
    if ((rc = W.wP1.ensure(N * 8))) return rc;
    if ((rc = W.wTmp.ensure((((u64)1 << dmax) >> W_B1) * 8))) return rc;   // one level-1 bucket
    if ((rc = W.wLlo.ensure(N * 4))) return rc;
    if ((rc = W.wLhi.ensure(N / 2))) return rc;
    if ((rc = W.wTab.ensure(2 * ((size_t)1 << W_AB) * 8))) return rc;
    // cursors: 256 level-1 buckets, then one per level-2 bucket of a
    // level-1 bucket (2^(d - 8 - 15) at the largest d)
    const size_t n2 = (size_t)1 << std::max(0, dmax - W_B1 - W_SB);
    if ((rc = W.wCur.ensure((256 + n2) * 4))) return rc;
    if ((rc = W.bHeavy.ensure((size_t)HEAVY_CAP * sizeof(Heavy)))) return rc;
    if ((rc = W.bCnt.ensure(64))) return rc;
    return GF2T_OK;
}

int sieve_wide(uint64_t r, int dmin, int dmax, uint64_t smax, uint64_t *dbest,
               gf2t_sieve_stats *stats) {
    SieveWork &W = *g_sw;
    int rc;
    if ((rc = wide_mem(dmax))) return rc;
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    const int place_lds = (int)((((size_t)1 << W_SB) + ((size_t)1 << W_SB) / 8) * sizeof(u32));
    std::call_once(lds_once, [place_lds] {
        lds_err = hipFuncSetAttribute((const void *)k_w_place,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, place_lds);
    });
    GF2T_HIP(lds_err);
    unsigned long long *cnt64 = (unsigned long long *)W.bCnt.p;
    unsigned *cnt32 = (unsigned *)((char *)W.bCnt.p + 32);
    unsigned *cur1 = (unsigned *)W.wCur.p, *cur2 = cur1 + 256;
    unsigned long long *P1 = (unsigned long long *)W.wP1.p, *tmp = (unsigned long long *)W.wTmp.p;
    u32 *Llo = (u32 *)W.wLlo.p, *Lhi = (u32 *)W.wLhi.p;
    u64 *A0 = (u64 *)W.wTab.p, *A1 = A0 + ((size_t)1 << W_AB);
    Heavy *heavy = (Heavy *)W.bHeavy.p;
    constexpr int MAX_SL = 16;
    hipEvent_t ev[4 + 2 * MAX_SL];
    for (auto &e : ev) GF2T_HIP(hipEventCreate(&e));
    struct Destroy {
        hipEvent_t *e;
        int n;
        ~Destroy() {
            for (int i = 0; i < n; i++) hipEventDestroy(e[i]);
        }
    } destroy{ev, 4 + 2 * MAX_SL};
    int status = GF2T_OK;
    for (int d = dmin; d <= dmax && status == GF2T_OK; d++) {
        const u64 M = ((u64)1 << d) - 1;
        const WideHost &H = wide_host(d);
        const u64 q = H.q, rmod = r % M;
        const int nsl = slices_for(d);
        if (d - W_B1 - W_SB < 2 || nsl > MAX_SL || 256 % nsl) {
            gf2t::set_error("gf2t_sieve: d=%d: the 64-bit path needs d >= 25 and at most %d "
                            "slices", d, MAX_SL);
            status = GF2T_EINVAL;
            break;
        }
        if (!q || gcd_host(rmod ? rmod : M, M) != 1) {
            gf2t::set_error("gf2t_sieve: d=%d: gcd(r, 2^d-1) != 1, unsupported r", d);
            status = GF2T_EINVAL;
            break;
        }
        if (!H.trace_ok) {
            gf2t::set_error("gf2t_sieve: d=%d: trace not in GF(2)", d);
            status = GF2T_EINTERNAL;
            break;
        }
        GF2T_HIP(hipMemcpy(A0, H.a0.data(), H.a0.size() * 8, hipMemcpyHostToDevice));
        GF2T_HIP(hipMemcpy(A1, H.a1.data(), H.a1.size() * 8, hipMemcpyHostToDevice));
        WParams P{d, M, q, rmod, smax, H.tmask, A0, A1};
        const u64 ncls = irreducible_count(d);
        const u64 qcap = ncls / 16 + 65536;
        if ((rc = W.wSurv.ensure((ncls + 64) * 8))) { status = rc; break; }
        if ((rc = W.wQueue.ensure(qcap * sizeof(MarkRecW)))) { status = rc; break; }
        u64 *surv = (u64 *)W.wSurv.p;
        MarkRecW *queue = (MarkRecW *)W.wQueue.p;

        // the classes (Lyndon words) first: every slice resolves its share
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, 0, cnt64, cnt32);
        GF2T_HIP(hipEventRecord(ev[0], 0));
        hipLaunchKernelGGL(k_lyndon_t<u64>, dim3(lyndon_blocks(d)), dim3(LY_T), 0, 0, d, M, surv,
                           cnt64, (u64)(ncls + 64));
        GF2T_CHECK_LAUNCH("k_lyndon_w");
        unsigned long long hs[2];
        GF2T_HIP(hipMemcpy(hs, cnt64, 16, hipMemcpyDeviceToHost));
        const u64 nsurv = hs[0];
        if (nsurv != ncls) {
            gf2t::set_error("gf2t_sieve: d=%d: %llu classes, expected %llu", d,
                            (unsigned long long)nsurv, (unsigned long long)ncls);
            status = GF2T_EINTERNAL;
            break;
        }
        GF2T_HIP(hipEventRecord(ev[1], 0));
        const int nb = 256 / nsl;
        const u64 cap1 = (u64)1 << (d - W_B1);                 // values per level-1 bucket
        const int B2 = d - W_B1 - W_SB, B2a = (B2 + 1) / 2, B2b = B2 - B2a;
        if ((1 << B2a) > WS_MAXB || W.wCur.bytes < (256 + ((size_t)1 << B2)) * 4) {
            gf2t::set_error("gf2t_sieve: d=%d: level-2 split of %d bits beyond the cursor or "
                            "digit tables", d, B2);
            status = GF2T_EINTERNAL;
            break;
        }
        const u64 capA = cap1 >> B2a;
        const u64 zb = std::min<u64>((nsurv + (u64)ZC_T * ZC_PER - 1) / ((u64)ZC_T * ZC_PER),
                                     1u << 16);
        for (int sl = 0; sl < nsl && status == GF2T_OK; sl++) {
            const unsigned b0 = (unsigned)(sl * nb);
            GF2T_HIP(hipEventRecord(ev[4 + 2 * sl], 0));
            // log table slice: partition pass 1 (this slice's buckets),
            // level-2 split per bucket, placement
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, 0, cur1, (u64)256);
            hipLaunchKernelGGL(k_w_part1, dim3((unsigned)((M + W_E - 1) / W_E)), dim3(W_T), 0, 0,
                               P, P1, cur1, cnt32 + 1, b0, (unsigned)nb);
            GF2T_CHECK_LAUNCH("k_w_part1");
            // level 2 per level-1 bucket: bucket -> scratch by the high half
            // of the level-2 bits, scratch -> bucket by the low half
            for (unsigned b1 = b0; b1 < b0 + (unsigned)nb; b1++) {
                const u64 count = b1 ? cap1 : cap1 - 1;   // value 0 never occurs
                unsigned long long *bk = P1 + (u64)(b1 - b0) * cap1;
                hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, 0, cur2, (u64)(1u << B2a));
                hipLaunchKernelGGL(k_w_split, dim3((unsigned)((count + W_E - 1) / W_E), 1),
                                   dim3(WS_T), 0, 0, bk, cap1, count, W_SB + B2b, B2a, tmp, capA,
                                   cur2, cnt32 + 1);
                hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)(((1u << B2) + 255) / 256)),
                                   dim3(256), 0, 0, cur2, (u64)(1u << B2));
                hipLaunchKernelGGL(k_w_split, dim3((unsigned)((capA + W_E - 1) / W_E), 1u << B2a),
                                   dim3(WS_T), 0, 0, tmp, capA, b1 ? capA : capA - 1, W_SB, B2b,
                                   bk, (u64)1 << W_SB, cur2, cnt32 + 1);
            }
            GF2T_CHECK_LAUNCH("k_w_split");
            hipLaunchKernelGGL(k_w_place, dim3((unsigned)(((u64)nb * cap1) >> W_SB)), dim3(1024),
                               (size_t)place_lds, 0, P1, Llo, Lhi, sl == 0 ? 1 : 0);
            GF2T_CHECK_LAUNCH("k_w_place");
            GF2T_HIP(hipEventRecord(ev[5 + 2 * sl], 0));
            unsigned chk[2];
            GF2T_HIP(hipMemcpy(chk, cnt32, 8, hipMemcpyDeviceToHost));
            if (chk[1]) {
                gf2t::set_error("gf2t_sieve: d=%d: log table partition overflow", d);
                status = GF2T_EINTERNAL;
                break;
            }
            hipLaunchKernelGGL(k_zech_w, dim3((unsigned)zb), dim3(ZC_T), 0, 0, P, surv, nsurv, Llo,
                               Lhi, queue, cnt64 + 1, qcap, (u64)b0 * cap1, (u64)nb * cap1);
            GF2T_CHECK_LAUNCH("k_zech_w");
        }
        if (status != GF2T_OK) break;
        GF2T_HIP(hipMemcpy(hs, cnt64, 16, hipMemcpyDeviceToHost));
        if (hs[1] > qcap) {
            gf2t::set_error("gf2t_sieve: d=%d: marking queue overflow (%llu)", d,
                            (unsigned long long)hs[1]);
            status = GF2T_EINTERNAL;
            break;
        }
        const u64 nq = hs[1];
        if (nq) {
            hipLaunchKernelGGL(k_mark_w, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, 0, P,
                               queue, nq, (unsigned long long *)dbest, heavy, cnt32, cnt32 + 1);
            GF2T_CHECK_LAUNCH("k_mark_w");
        }
        GF2T_HIP(hipEventRecord(ev[2], 0));
        unsigned h32[2];
        GF2T_HIP(hipMemcpy(h32, cnt32, 8, hipMemcpyDeviceToHost));
        if (h32[1]) {
            gf2t::set_error("gf2t_sieve: d=%d: %s", d,
                            (h32[1] & 1) ? "minimal polynomial not over GF(2)"
                                         : "heavy queue overflow");
            status = GF2T_EINTERNAL;
            break;
        }
        const unsigned nheavy = h32[0];
        if (nheavy) {
            unsigned gy = nheavy < 65535 ? nheavy : 65535;
            hipLaunchKernelGGL(k_mark_heavy, dim3(64, gy), dim3(256), 0, 0, heavy, nheavy, smax,
                               (unsigned long long *)dbest);
            GF2T_CHECK_LAUNCH("k_mark_heavy");
        }
        GF2T_HIP(hipEventRecord(ev[3], 0));
        GF2T_HIP(hipEventSynchronize(ev[3]));
        if (stats) {
            gf2t_sieve_stats &S = stats[d];
            S.d = d;
            S.q = q;
            S.classes = nsurv;
            S.marking = nq;
            S.heavy = nheavy;
            // table: the slices' builds; classes: the rest up to marking
            float tab = 0, all = 0, x;
            for (int sl = 0; sl < nsl; sl++) {
                GF2T_HIP(hipEventElapsedTime(&x, ev[4 + 2 * sl], ev[5 + 2 * sl]));
                tab += x;
            }
            GF2T_HIP(hipEventElapsedTime(&all, ev[0], ev[2]));
            S.ms_table = tab;
            S.ms_classes = all - tab;
            GF2T_HIP(hipEventElapsedTime(&S.ms_heavy, ev[2], ev[3]));
        }
    }
    return status;
}
}  // namespace

// gf2t_tune keys of the sieve (test knobs): "sieve_wide_from" = first
// degree of the 64-bit path (25..33, default 33), "sieve_slices" = least
// number of log-table slices (1..16, a power of two)
int gf2t_sieve_tune(const char *key, int64_t v) {
    std::lock_guard<std::mutex> lk(g_sw_mu);
    if (!strcmp(key, "sieve_wide_from") && v >= 25 && v <= 33) {
        g_wide_from = (int)v;
        return GF2T_OK;
    }
    if (!strcmp(key, "sieve_slices") && v >= 1 && v <= 16 && !(v & (v - 1))) {
        g_min_slices = (int)v;
        return GF2T_OK;
    }
    gf2t::set_error("gf2t_tune: %s needs 25..33 (sieve_wide_from) or a power of two 1..16 "
                    "(sieve_slices)", key);
    return GF2T_EINVAL;
}
int gf2t_sieve_tune_get(const char *key, int64_t *v) {
    std::lock_guard<std::mutex> lk(g_sw_mu);
    if (!strcmp(key, "sieve_wide_from")) { *v = g_wide_from; return GF2T_OK; }
    if (!strcmp(key, "sieve_slices")) { *v = g_min_slices; return GF2T_OK; }
    return GF2T_EINVAL;
}

extern "C" int gf2t_sieve(uint64_t r, int dmin, int dmax, uint64_t smax, uint64_t *dbest,
                          gf2t_sieve_stats *stats) {
    if (dmin < 2) dmin = 2;
    if (dmax > WIDE_MAX_D || dmin > dmax || !dbest || smax < 1) {
        gf2t::set_error("gf2t_sieve: need 2 <= dmin <= dmax <= %d, smax >= 1, dbest",
                        WIDE_MAX_D);
        return GF2T_EINVAL;
    }
    // device tables stay allocated between calls (grow-only; freed by
    // gf2t_sieve_release), so repeated sieves pay no allocation
    std::lock_guard<std::mutex> lk(g_sw_mu);
    if (!g_sw) g_sw = new SieveWork();
    int status = GF2T_OK;
    const int wf = g_wide_from;
    if (dmax >= 33) {
        // the degree-<=32 arrays live inside the wide tables from here on
        for (DevBuf *b : {&g_sw->bL, &g_sw->bP1, &g_sw->bP2, &g_sw->bSurv, &g_sw->bQueue})
            b->release();
        status = wide_mem(dmax);
    }
    if (status == GF2T_OK && dmin < wf)
        status = sieve_narrow(r, dmin, std::min(dmax, wf - 1), smax, dbest, stats);
    if (status == GF2T_OK && dmax >= wf)
        status = sieve_wide(r, std::max(dmin, wf), dmax, smax, dbest, stats);
    return status;
}
