/*
 * gf2_oracle.c -- CPU restatement of the reference algorithms.
 * TEST INFRASTRUCTURE ONLY: see gf2_oracle.h for what pins each part.
 * Reference citations are file:line in the reference tree.
 */
#define _GNU_SOURCE
#include "gf2_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <wmmintrin.h>
#include <smmintrin.h>

#define WORDS(bits) ((size_t)(((bits) + 63) >> 6))

/* ======================================================================
 * GF(2^d), d <= 32.  Elements are u32 bit masks, modulus q has bit d set
 * (reference: src/gf2_small_field.h:30-111).
 * ==================================================================== */
uint64_t gf2o_clmul32(uint32_t a, uint32_t b) {
    __m128i p = _mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a),
                                     _mm_cvtsi64_si128((long long)b), 0);
    return (uint64_t)_mm_cvtsi128_si64(p);
}

static uint32_t fred(uint64_t t, uint64_t q, int d) {
    /* polynomial remainder of t (degree <= 2d-2) by q, top bit first */
    for (int b = 2 * d - 2; b >= d; b--)
        if ((t >> b) & 1) t ^= q << (b - d);
    return (uint32_t)t;
}

uint32_t gf2o_fmul(uint32_t a, uint32_t b, uint64_t q, int d) {
    return fred(gf2o_clmul32(a, b), q, d);
}

static uint32_t fmulx(uint32_t a, uint64_t q, int d) {
    uint64_t t = (uint64_t)a << 1;
    if ((t >> d) & 1) t ^= q;
    return (uint32_t)t;
}

static uint32_t fpow(uint32_t base, uint64_t e, uint64_t q, int d) {
    uint32_t r = 1;
    while (e) {
        if (e & 1) r = gf2o_fmul(r, base, q, d);
        base = gf2o_fmul(base, base, q, d);
        e >>= 1;
    }
    return r;
}

static int prime_factors(uint64_t m, uint64_t *ps) {
    int n = 0;
    for (uint64_t p = 2; p * p <= m; p += (p == 2 ? 1 : 2))
        if (m % p == 0) { ps[n++] = p; while (m % p == 0) m /= p; }
    if (m > 1) ps[n++] = m;
    return n;
}

/* 64-bit-coefficient helpers for polynomials of degree <= 62 */
static uint64_t pmod64(uint64_t a, uint64_t q, int d) {
    for (int b = 63; b >= d; b--)
        if ((a >> b) & 1) a ^= q << (b - d);
    return a;
}
static uint64_t pmulmod64(uint64_t a, uint64_t b, uint64_t q, int d) {
    /* a, b < 2^d; full 128-bit carry-less product then remainder */
    __m128i p = _mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a),
                                     _mm_cvtsi64_si128((long long)b), 0);
    uint64_t lo = (uint64_t)_mm_cvtsi128_si64(p);
    uint64_t hi = (uint64_t)_mm_extract_epi64(p, 1);
    for (int bit = 2 * d - 2; bit >= d; bit--) {
        uint64_t have = bit >= 64 ? (hi >> (bit - 64)) & 1 : (lo >> bit) & 1;
        if (!have) continue;
        int sh = bit - d;
        lo ^= q << sh;
        if (sh) hi ^= q >> (64 - sh);
    }
    return lo;
}
static uint64_t pgcd64(uint64_t a, uint64_t b) {
    while (b) {
        int db = 63 - __builtin_clzll(b);
        while (a && 63 - __builtin_clzll(a) >= db)
            a ^= b << ((63 - __builtin_clzll(a)) - db);
        uint64_t t = a; a = b; b = t;
    }
    return a;
}

/* Rabin's test: x^(2^d) == x mod q and gcd(x^(2^(d/p)) - x, q) == 1
 * for every prime p | d.  Independent of the reference's variant
 * (src/gf2_field_setup.h:25-35), which omits the gcd. */
int gf2o_is_irreducible_small(uint64_t q, int d) {
    if (d < 1 || d > 62 || !((q >> d) & 1)) return 0;
    if (q >> (d + 1)) return 0;
    if (d == 1) return 1;
    uint64_t chain[64];
    uint64_t t = 2;
    chain[0] = t;
    for (int i = 1; i <= d; i++) { t = pmulmod64(t, t, q, d); chain[i] = t; }
    if (chain[d] != 2) return 0;
    uint64_t ps[16];
    int np = prime_factors((uint64_t)d, ps);
    for (int i = 0; i < np; i++) {
        uint64_t h = chain[d / ps[i]] ^ 2;
        if (pgcd64(q, h) != 1) return 0;
    }
    return 1;
}

int gf2o_is_irreducible(uint64_t q, int d) { return gf2o_is_irreducible_small(q, d); }

uint64_t gf2o_find_primitive_poly(int d) {
    /* smallest primitive polynomial of degree d (reference:
     * src/gf2_field_setup.h:49-57); the sieve output does not depend on
     * which primitive modulus is used, this choice only keeps tables
     * comparable with the reference's. */
    uint64_t M = ((uint64_t)1 << d) - 1, ps[16];
    int np = prime_factors(M, ps);
    for (uint64_t q = ((uint64_t)1 << d) | 1; q < ((uint64_t)2 << d); q += 2) {
        if (!gf2o_is_irreducible_small(q, d)) continue;
        int prim = 1;
        for (int i = 0; i < np && prim; i++)
            if (fpow(2, M / ps[i], q, d) == 1) prim = 0;
        if (prim) return q;
    }
    return 0;
}

static int64_t mobius(uint64_t n) {
    int64_t mu = 1;
    for (uint64_t p = 2; p * p <= n; p++) {
        if (n % p) continue;
        n /= p;
        if (n % p == 0) return 0;
        mu = -mu;
    }
    if (n > 1) mu = -mu;
    return mu;
}

uint64_t gf2o_count_irreducible(int d) {
    /* (1/d) sum_{k | d} mu(k) 2^(d/k) */
    int64_t acc = 0;
    for (int k = 1; k <= d; k++)
        if (d % k == 0) acc += mobius((uint64_t)k) * ((int64_t)1 << (d / k));
    return (uint64_t)(acc / d);
}

/* ======================================================================
 * Coarse sieve (reference: src/sieve_ref.cpp:58-123, the algorithm the
 * GPU sieve src/coarse_sieve.cu:237-288 also follows).
 * ==================================================================== */
static uint64_t gcd_u(uint64_t a, uint64_t b) {
    while (b) { uint64_t t = a % b; a = b; b = t; }
    return a;
}
static uint64_t inv_u(uint64_t a, uint64_t m) {
    int64_t t = 0, nt = 1, rr = (int64_t)m, nr = (int64_t)(a % m);
    while (nr) {
        int64_t qq = rr / nr, tmp;
        tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = rr - qq * nr; rr = nr; nr = tmp;
    }
    if (t < 0) t += (int64_t)m;
    return (uint64_t)t;
}

static inline void amin64(uint64_t *p, uint64_t v) __attribute__((unused));
static inline void amin64(uint64_t *p, uint64_t v) {
    uint64_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur &&
           !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED)) {}
}

/* one marking progression: s = sf, sf + n, ... (<= smax) gets key */
struct prog { uint64_t sf, n, key; };

struct sieve_job {
    int d; uint64_t M, q, rmod, smax, j0, j1;   /* j0: first block, j1: block stride */
    const uint32_t *A, *L;
    uint64_t *best;
    uint64_t classes;
    struct prog *pg;          /* progressions found by this job */
    size_t npg, cap;
    uint64_t lo, hi;          /* marking phase: this job's s range [lo, hi) */
    struct sieve_job *all;    /* every job's progressions (marking phase) */
    int njobs;
};

static void *sieve_worker(void *arg) {
    struct sieve_job *J = (struct sieve_job *)arg;
    const int d = J->d;
    const uint64_t M = J->M, q = J->q;
    uint32_t P[40];
    /* blocks of 4096 candidates dealt round-robin: the classes that
     * survive (Lyndon words) all lie below 2^(d-1), so contiguous ranges
     * would leave the upper threads idle */
    for (uint64_t b0 = J->j0; b0 < J->M; b0 += J->j1)
    for (uint64_t j = b0 ? b0 : 1; j < b0 + 4096 && j < J->M; j++) {
        /* keep j iff it is the strict minimum of its rotation orbit
         * {j*2^i mod M}: that excludes proper-subfield elements (orbit
         * shorter than d) and non-minimal conjugates alike
         * (reference: src/sieve_ref.cpp:87-98) */
        uint64_t jj = j;
        int keep = 1;
        for (int i = 1; i < d; i++) {
            jj = ((jj << 1) | (jj >> (d - 1))) & M;
            if (jj <= j) { keep = 0; break; }
        }
        if (!keep) continue;
        J->classes++;
        uint64_t u = (j * J->rmod) % M;
        if (u == 0) continue;
        uint64_t v = J->L[J->A[u] ^ 1u];          /* Zech logarithm */
        uint64_t gg = gcd_u(j, M);
        if (v % gg) continue;
        uint64_t n = M / gg;
        uint64_t s0 = (inv_u(j / gg, n) * (v / gg)) % n;
        uint64_t sf = s0 ? s0 : n;
        if (sf > J->smax) continue;
        /* minimal polynomial prod_i (X + g^(j 2^i)) (reference:
         * src/gf2_small_field.h:117-135) */
        for (int k = 0; k <= d; k++) P[k] = 0;
        P[0] = 1;
        uint64_t e = j;
        for (int i = 0; i < d; i++) {
            uint32_t root = J->A[e];
            for (int k = i + 1; k >= 1; k--)
                P[k] = P[k - 1] ^ gf2o_fmul(root, P[k], q, d);
            P[0] = gf2o_fmul(root, P[0], q, d);
            e = ((e << 1) | (e >> (d - 1))) & M;
        }
        uint64_t mask = 0;
        for (int k = 0; k <= d; k++) {
            if (P[k] > 1) return (void *)1;       /* not in GF(2): bug */
            mask |= (uint64_t)P[k] << k;
        }
        if (J->npg == J->cap) {
            J->cap = J->cap ? 2 * J->cap : 1024;
            struct prog *np = (struct prog *)realloc(J->pg, J->cap * sizeof *np);
            if (!np) return (void *)1;
            J->pg = np;
        }
        J->pg[J->npg++] = (struct prog){sf, n, ((uint64_t)d << 48) | mask};
    }
    return 0;
}

/* marking phase: every progression of every job, restricted to this
 * job's s range, so no two threads write the same best[s] */
static void *mark_worker(void *arg) {
    struct sieve_job *J = (struct sieve_job *)arg;
    for (int k = 0; k < J->njobs; k++) {
        const struct sieve_job *O = &J->all[k];
        for (size_t p = 0; p < O->npg; p++) {
            const struct prog g = O->pg[p];
            uint64_t s = g.sf;
            if (s < J->lo) s += (J->lo - s + g.n - 1) / g.n * g.n;
            for (; s < J->hi; s += g.n)
                if (g.key < J->best[s]) J->best[s] = g.key;
        }
    }
    return 0;
}

/* Antilog and log tables of GF(2^d) = GF(2)[x]/(q) for i in [i0, i1):
 * A[i] = x^i, L[x^i] = i (the table build of src/sieve_ref.cpp, split
 * into ranges so threads can share it; each range starts at x^i0). */
struct table_job {
    int d; uint64_t q, i0, i1;
    uint32_t *A, *L;
};

static void *table_worker(void *arg) {
    struct table_job *T = (struct table_job *)arg;
    uint32_t t = fpow(2u, T->i0, T->q, T->d);
    for (uint64_t i = T->i0; i < T->i1; i++) {
        T->A[i] = t;
        T->L[t] = (uint32_t)i;
        t = fmulx(t, T->q, T->d);
    }
    return 0;
}

int gf2o_sieve(uint64_t r, int dmin, int dmax, uint64_t smax, uint64_t *best,
               uint64_t *classes, int nthreads) {
    if (dmin < 2) dmin = 2;
    if (dmax > 32 || nthreads < 1) return -1;
    for (int d = dmin; d <= dmax; d++) {
        const uint64_t M = ((uint64_t)1 << d) - 1;
        const uint64_t q = gf2o_find_primitive_poly(d);
        if (!q) return -2;
        const uint64_t rmod = r % M;
        if (gcd_u(rmod ? rmod : M, M) != 1) return -3;
        uint32_t *A = (uint32_t *)malloc(M * sizeof(uint32_t));
        uint32_t *L = (uint32_t *)malloc((M + 1) * sizeof(uint32_t));
        if (!A || !L) { free(A); free(L); return -4; }
        int nt = nthreads;
        if ((uint64_t)nt > M) nt = 1;
        {
            struct table_job *tj = (struct table_job *)calloc(nt, sizeof *tj);
            pthread_t *tt = (pthread_t *)calloc(nt, sizeof *tt);
            for (int k = 0; k < nt; k++) {
                tj[k] = (struct table_job){d, q, M * k / nt, M * (k + 1) / nt, A, L};
                if (nt == 1) table_worker(&tj[k]);
                else pthread_create(&tt[k], 0, table_worker, &tj[k]);
            }
            if (nt > 1)
                for (int k = 0; k < nt; k++) pthread_join(tt[k], 0);
            free(tj);
            free(tt);
        }
        if (fmulx(A[M - 1], q, d) != 1) { free(A); free(L); return -5; }
        struct sieve_job *jobs = (struct sieve_job *)calloc(nt, sizeof *jobs);
        pthread_t *th = (pthread_t *)calloc(nt, sizeof *th);
        int bad = 0;
        for (int k = 0; k < nt; k++) {
            struct sieve_job *J = &jobs[k];
            J->d = d; J->M = M; J->q = q; J->rmod = rmod; J->smax = smax;
            J->j0 = 4096 * (uint64_t)k;        /* first block of this thread */
            J->j1 = 4096 * (uint64_t)nt;       /* stride between its blocks */
            J->A = A; J->L = L; J->best = best;
            if (nt == 1) { bad |= sieve_worker(J) != 0; }
            else pthread_create(&th[k], 0, sieve_worker, J);
        }
        uint64_t cls = 0;
        for (int k = 0; k < nt; k++) {
            if (nt > 1) { void *rv; pthread_join(th[k], &rv); bad |= rv != 0; }
            cls += jobs[k].classes;
        }
        if (classes) classes[d] = cls;
        if (!bad) {   /* mark: the s range split among the threads */
            for (int k = 0; k < nt; k++) {
                struct sieve_job *J = &jobs[k];
                J->lo = 1 + smax * (uint64_t)k / nt;
                J->hi = 1 + smax * (uint64_t)(k + 1) / nt;
                J->all = jobs;
                J->njobs = nt;
                if (nt == 1) mark_worker(J);
                else pthread_create(&th[k], 0, mark_worker, J);
            }
            if (nt > 1)
                for (int k = 0; k < nt; k++) pthread_join(th[k], 0);
        }
        for (int k = 0; k < nt; k++) free(jobs[k].pg);
        free(jobs); free(th); free(A); free(L);
        if (bad) return -6;
    }
    return 0;
}

/* Brute force: every irreducible P of degree 2..depth, x^r + x^s + 1
 * evaluated mod P with x^s stepped incrementally (reference:
 * src/sieve_ref.cpp:142-160). */
int gf2o_brute_force(uint64_t r, int depth, uint64_t smax, uint64_t *best) {
    for (int d = 2; d <= depth; d++) {
        for (uint64_t q = ((uint64_t)1 << d) | 1; q < ((uint64_t)2 << d); q += 2) {
            if (!gf2o_is_irreducible_small(q, d)) continue;
            uint64_t xr = 1, b = 2, e = r;
            while (e) {
                if (e & 1) xr = pmulmod64(xr, b, q, d);
                b = pmulmod64(b, b, q, d);
                e >>= 1;
            }
            uint64_t xs = 1, key = ((uint64_t)d << 48) | q;
            for (uint64_t s = 1; s <= smax; s++) {
                xs <<= 1;
                if ((xs >> d) & 1) xs ^= q;
                if ((xr ^ xs ^ 1) == 0 && key < best[s]) best[s] = key;
            }
        }
    }
    return 0;
}

int gf2o_divides_trinomial_small(uint64_t r, uint64_t s, uint64_t q, int d) {
    if (d < 1 || d > 62) return 0;
    uint64_t acc[2] = {1, 1}, b = pmod64(2, q, d), es[2] = {r, s};
    for (int k = 0; k < 2; k++) {
        uint64_t bb = b, e = es[k];
        while (e) {
            if (e & 1) acc[k] = pmulmod64(acc[k], bb, q, d);
            bb = pmulmod64(bb, bb, q, d);
            e >>= 1;
        }
    }
    return (acc[0] ^ acc[1] ^ pmod64(1, q, d)) == 0;
}

/* ======================================================================
 * Residues mod T = x^r + x^s + 1, 1 <= s <= (r-1)/2.
 * ==================================================================== */
static uint64_t spread32(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFULL;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFULL;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0FULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

/* dst ^= src << sh, src has `sbits` significant bits; dst big enough */
static void xor_shl(uint64_t *dst, const uint64_t *src, uint64_t sbits,
                    uint64_t sh) {
    size_t sw = WORDS(sbits), ws = (size_t)(sh >> 6);
    unsigned bs = (unsigned)(sh & 63);
    for (size_t j = 0; j < sw; j++) {
        uint64_t v = src[j];
        dst[j + ws] ^= v << bs;
        if (bs && (v >> (64 - bs))) dst[j + ws + 1] ^= v >> (64 - bs);
    }
}

/* Two-pass trinomial fold x^e -> x^(e-r) + x^(e-r+s) (the reference's
 * 3-kernel squaring path, src/gf2_trinomial_cuda.h:44-74, restated at
 * word level): the part above bit r is split off and added back at
 * offsets 0 and s; at most two rounds since 2s < r. */
void gf2o_reduce_fold(uint64_t *v, uint64_t bits, uint64_t r, uint64_t s) {
    while (bits > r) {
        uint64_t hb = bits - r;
        size_t hw = WORDS(hb), w0 = (size_t)(r >> 6);
        unsigned sh = (unsigned)(r & 63);
        uint64_t *H = (uint64_t *)calloc(hw + 1, 8);
        size_t vw = WORDS(bits);
        for (size_t k = 0; k < hw; k++) {
            uint64_t a = v[w0 + k];
            uint64_t b = (w0 + k + 1 < vw) ? v[w0 + k + 1] : 0;
            H[k] = sh ? ((a >> sh) | (b << (64 - sh))) : a;
        }
        if (hb & 63) H[hw - 1] &= (~0ULL) >> (64 - (hb & 63));
        /* clear bits >= r */
        if (sh) v[w0] &= (1ULL << sh) - 1; else v[w0] = 0;
        for (size_t k = w0 + 1; k < vw; k++) v[k] = 0;
        xor_shl(v, H, hb, 0);
        xor_shl(v, H, hb, s);
        free(H);
        uint64_t nb = hb + s;
        bits = nb > r ? nb : r;
    }
}

/* bit-serial reduction from the top (reference:
 * src/gf2_trinomial_cuda.h:182-190) */
void gf2o_reduce_bitserial(uint64_t *v, uint64_t bits, uint64_t r, uint64_t s) {
    for (uint64_t e = bits; e-- > r;) {
        if (!((v[e >> 6] >> (e & 63)) & 1)) continue;
        v[e >> 6] ^= 1ULL << (e & 63);
        uint64_t e1 = e - r, e2 = e - r + s;
        v[e1 >> 6] ^= 1ULL << (e1 & 63);
        v[e2 >> 6] ^= 1ULL << (e2 & 63);
    }
}

void gf2o_square_fold(const uint64_t *in, uint64_t *out, uint64_t r, uint64_t s) {
    size_t nw = WORDS(r);
    uint64_t *S = (uint64_t *)calloc(2 * nw + 2, 8);
    for (size_t i = 0; i < nw; i++) {
        S[2 * i] = spread32((uint32_t)in[i]);
        S[2 * i + 1] = spread32((uint32_t)(in[i] >> 32));
    }
    gf2o_reduce_fold(S, 2 * r - 1, r, s);
    memcpy(out, S, nw * 8);
    free(S);
}

void gf2o_square_fold_steps(uint64_t *res, uint64_t r, uint64_t s, uint64_t steps) {
    size_t nw = WORDS(r);
    uint64_t *S = (uint64_t *)calloc(2 * nw + 2, 8);
    for (uint64_t k = 0; k < steps; k++) {
        memset(S, 0, (2 * nw + 2) * 8);
        for (size_t i = 0; i < nw; i++) {
            S[2 * i] = spread32((uint32_t)res[i]);
            S[2 * i + 1] = spread32((uint32_t)(res[i] >> 32));
        }
        gf2o_reduce_fold(S, 2 * r - 1, r, s);
        memcpy(res, S, nw * 8);
    }
    free(S);
}

void gf2o_mulx_mod(const uint64_t *in, uint64_t *out, uint64_t r, uint64_t s) {
    size_t nw = WORDS(r);
    uint64_t carry = 0;
    for (size_t i = 0; i < nw; i++) {
        uint64_t v = in[i];
        out[i] = (v << 1) | carry;
        carry = v >> 63;
    }
    uint64_t top = (r >> 6) < nw ? (out[r >> 6] >> (r & 63)) & 1 : carry;
    if ((r >> 6) < nw) out[r >> 6] &= ~(1ULL << (r & 63));
    if (top) { out[0] ^= 1; out[s >> 6] ^= 1ULL << (s & 63); }
}

/* ======================================================================
 * GF(2)[x] products
 * ==================================================================== */
size_t gf2o_mul_out_words(uint64_t abits, uint64_t bbits) {
    return WORDS(abits + bbits) + 1;
}

void gf2o_mul_naive(const uint64_t *a, uint64_t abits, const uint64_t *b,
                    uint64_t bbits, uint64_t *out) {
    size_t ow = gf2o_mul_out_words(abits, bbits), wb = WORDS(bbits);
    memset(out, 0, ow * 8);
    for (uint64_t i = 0; i < abits; i++) {
        if (!((a[i >> 6] >> (i & 63)) & 1)) continue;
        xor_shl(out, b, (uint64_t)wb * 64, i);
    }
}

static void clmul64(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi) {
    __m128i p = _mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a),
                                     _mm_cvtsi64_si128((long long)b), 0);
    *lo = (uint64_t)_mm_cvtsi128_si64(p);
    *hi = (uint64_t)_mm_extract_epi64(p, 1);
}

void gf2o_mul_comb(const uint64_t *a, uint64_t abits, const uint64_t *b,
                   uint64_t bbits, uint64_t *out) {
    size_t wa = WORDS(abits), wb = WORDS(bbits);
    size_t ow = gf2o_mul_out_words(abits, bbits);
    memset(out, 0, ow * 8);
    for (size_t i = 0; i < wa; i++)
        for (size_t j = 0; j < wb; j++) {
            uint64_t lo, hi;
            clmul64(a[i], b[j], &lo, &hi);
            out[i + j] ^= lo;
            out[i + j + 1] ^= hi;
        }
}

/* ======================================================================
 * GF(2^64) = GF(2)[x]/(x^64 + x^4 + x^3 + x + 1) and the Cantor FFT
 * ==================================================================== */
uint64_t gf2o_gf64_mul_bitserial(uint64_t a, uint64_t b) {
    uint64_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        uint64_t hi = a >> 63;
        a <<= 1;
        if (hi) a ^= 0x1BULL;
        b >>= 1;
    }
    return r;
}

uint64_t gf2o_gf64_mul(uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    clmul64(a, b, &lo, &hi);
    /* x^64 = x^4 + x^3 + x + 1: fold hi twice */
    uint64_t l2, h2;
    clmul64(hi, 0x1BULL, &l2, &h2);
    lo ^= l2;
    uint64_t l3, h3;
    clmul64(h2, 0x1BULL, &l3, &h3);
    return lo ^ l3;
}

/* Cantor basis b[0] = 1, b[i]^2 + b[i] = b[i-1].  y^2 + y = c has the two
 * roots y and y+1 (the kernel of T(y) = y^2 + y is {0,1}); the reference
 * (src/gf2_cantor_engine.h:78-110) returns the root with bit 0 clear,
 * since its elimination never uses the kernel vector.  Solved here by
 * plain Gauss-Jordan elimination on the 64x64 matrix of T. */
static int solve_T(uint64_t c, uint64_t *y) {
    uint64_t row[64], rhs[64];
    /* row j: coefficient mask over unknowns y_i of output bit j */
    for (int j = 0; j < 64; j++) { row[j] = 0; rhs[j] = (c >> j) & 1; }
    for (int i = 0; i < 64; i++) {
        uint64_t ei = 1ULL << i;
        uint64_t t = gf2o_gf64_mul(ei, ei) ^ ei;
        for (int j = 0; j < 64; j++)
            if ((t >> j) & 1) row[j] |= ei;
    }
    int piv_of_col[64];
    for (int i = 0; i < 64; i++) piv_of_col[i] = -1;
    int rank = 0;
    for (int col = 63; col >= 0; col--) {
        int p = -1;
        for (int j = rank; j < 64; j++) if ((row[j] >> col) & 1) { p = j; break; }
        if (p < 0) continue;
        uint64_t tr = row[p]; row[p] = row[rank]; row[rank] = tr;
        uint64_t tb = rhs[p]; rhs[p] = rhs[rank]; rhs[rank] = tb;
        for (int j = 0; j < 64; j++)
            if (j != rank && ((row[j] >> col) & 1)) { row[j] ^= row[rank]; rhs[j] ^= rhs[rank]; }
        piv_of_col[col] = rank++;
    }
    for (int j = rank; j < 64; j++) if (rhs[j]) return 0;     /* Tr(c) = 1 */
    uint64_t sol = 0;                                         /* free vars = 0 */
    for (int col = 0; col < 64; col++)
        if (piv_of_col[col] >= 0 && rhs[piv_of_col[col]]) sol |= 1ULL << col;
    if (sol & 1) sol ^= 1;                                    /* bit 0 clear */
    if ((gf2o_gf64_mul(sol, sol) ^ sol) != c) return 0;
    *y = sol;
    return 1;
}

int gf2o_cantor_basis(uint64_t b[64]) {
    b[0] = 1;
    int len = 1;
    while (len < 64) {
        uint64_t y;
        if (!solve_T(b[len - 1], &y)) break;
        b[len++] = y;
    }
    return len;
}

/* point of index idx: sum of b[k] over the set bits k of idx */
void gf2o_cantor_point(const uint64_t *basis, uint64_t idx, uint64_t *pt) {
    uint64_t p = 0;
    for (int k = 0; k < 64; k++) if ((idx >> k) & 1) p ^= basis[k];
    *pt = p;
}

uint64_t gf2o_poly_eval64(const uint64_t *f, size_t n, uint64_t pt) {
    uint64_t acc = 0;
    for (size_t k = n; k-- > 0;) acc = gf2o_gf64_mul(acc, pt) ^ f[k];
    return acc;
}

/* Taylor level e: on index bits (e+1, e), [10] ^= [11] then [01] ^= [10].
 * The reference's range form (src/gf2_cantor_engine.h:149-158) applies
 * exactly these two XOR sweeps for block size B = 2^(e-L+2) at depth L. */
static void taylor_level(uint64_t *f, size_t n, int e, int inverse) {
    size_t st = (size_t)1 << e, blk = st << 2;
    for (size_t o = 0; o < n; o += blk) {
        uint64_t *p01 = f + o + st, *p10 = f + o + 2 * st, *p11 = f + o + 3 * st;
        if (!inverse) {
            for (size_t k = 0; k < st; k++) p10[k] ^= p11[k];
            for (size_t k = 0; k < st; k++) p01[k] ^= p10[k];
        } else {
            for (size_t k = 0; k < st; k++) p01[k] ^= p10[k];
            for (size_t k = 0; k < st; k++) p10[k] ^= p11[k];
        }
    }
}

/* twiddle of pair-group t: sum of b[k+1] over set bits k of t
 * (reference: src/gf2_cantor_engine.h:130-139) */
static uint64_t *twiddles(const uint64_t *basis, int m) {
    size_t h = (size_t)1 << (m > 0 ? m - 1 : 0);
    uint64_t *W = (uint64_t *)calloc(h, 8);
    for (size_t t = 1; t < h; t++) {
        int hb = 63 - __builtin_clzll((uint64_t)t);
        W[t] = W[t ^ ((size_t)1 << hb)] ^ basis[hb + 1];
    }
    return W;
}

static void bfly_level(uint64_t *f, size_t n, int L, const uint64_t *W, int inverse) {
    size_t st = (size_t)1 << L;
    for (size_t i1 = 0; i1 < n; i1++) {
        if (i1 & st) continue;
        size_t i2 = i1 | st;
        uint64_t w = W[i1 >> (L + 1)];
        if (!inverse) {
            uint64_t e1 = f[i1] ^ gf2o_gf64_mul(w, f[i2]);
            f[i1] = e1;
            f[i2] = e1 ^ f[i2];
        } else {
            uint64_t h = f[i1] ^ f[i2];
            f[i1] = f[i1] ^ gf2o_gf64_mul(w, h);
            f[i2] = h;
        }
    }
}

void gf2o_fft_fwd(uint64_t *f, int m, const uint64_t *basis) {
    size_t n = (size_t)1 << m;
    for (int L = 0; L <= m - 2; L++)
        for (int e = m - 2; e >= L; e--) taylor_level(f, n, e, 0);
    uint64_t *W = twiddles(basis, m);
    for (int L = m - 1; L >= 0; L--) bfly_level(f, n, L, W, 0);
    free(W);
}

void gf2o_fft_inv(uint64_t *f, int m, const uint64_t *basis) {
    size_t n = (size_t)1 << m;
    uint64_t *W = twiddles(basis, m);
    for (int L = 0; L <= m - 1; L++) bfly_level(f, n, L, W, 1);
    free(W);
    for (int L = m - 2; L >= 0; L--)
        for (int e = L; e <= m - 2; e++) taylor_level(f, n, e, 1);
}

/* Product by the Cantor transform (reference:
 * src/gf2_cantor_engine.h:192-242): 32-bit chunks lifted to GF(2^64),
 * 2^m >= ca + cb - 1 points, overlap-add at 32-bit stride. */
int gf2o_fft_mul(const uint64_t *a, uint64_t abits, const uint64_t *b,
                 uint64_t bbits, uint64_t *out) {
    uint64_t basis[64];
    if (gf2o_cantor_basis(basis) != 64) return -1;
    size_t ca = (size_t)((abits + 31) >> 5), cb = (size_t)((bbits + 31) >> 5);
    if (!ca || !cb) { memset(out, 0, gf2o_mul_out_words(abits, bbits) * 8); return 0; }
    int m = 1;
    while (((size_t)1 << m) < ca + cb - 1) m++;
    size_t n = (size_t)1 << m;
    uint64_t *FA = (uint64_t *)calloc(n, 8), *FB = (uint64_t *)calloc(n, 8);
    if (!FA || !FB) { free(FA); free(FB); return -2; }
    const uint32_t *a32 = (const uint32_t *)a, *b32 = (const uint32_t *)b;
    for (size_t k = 0; k < ca; k++) FA[k] = a32[k];
    for (size_t k = 0; k < cb; k++) FB[k] = b32[k];
    if (abits & 31) FA[ca - 1] &= (1ULL << (abits & 31)) - 1;
    if (bbits & 31) FB[cb - 1] &= (1ULL << (bbits & 31)) - 1;
    gf2o_fft_fwd(FA, m, basis);
    gf2o_fft_fwd(FB, m, basis);
    for (size_t k = 0; k < n; k++) FA[k] = gf2o_gf64_mul(FA[k], FB[k]);
    gf2o_fft_inv(FA, m, basis);
    size_t ow = gf2o_mul_out_words(abits, bbits), nc = ca + cb - 1;
    for (size_t w = 0; w < ow; w++) {
        uint64_t v = 0;
        size_t k = 2 * w;
        if (k < nc) v ^= FA[k];
        if (k + 1 < nc) v ^= FA[k + 1] << 32;
        if (w && k - 1 < nc) v ^= FA[k - 1] >> 32;
        out[w] = v;
    }
    free(FA); free(FB);
    return 0;
}

int gf2o_mulmod(const uint64_t *a, uint64_t abits, const uint64_t *b,
                uint64_t bbits, uint64_t r, uint64_t s, uint64_t *out) {
    size_t ow = gf2o_mul_out_words(abits, bbits), nw = WORDS(r);
    size_t cap = ow > 2 * nw + 2 ? ow : 2 * nw + 2;
    uint64_t *P = (uint64_t *)calloc(cap, 8);
    int rc = gf2o_fft_mul(a, abits, b, bbits, P);
    if (rc) { free(P); return rc; }
    uint64_t pbits = abits + bbits - 1;          /* bits of the product */
    if (pbits > 2 * r - 1) { free(P); return -3; }
    gf2o_reduce_fold(P, pbits < r ? r : pbits, r, s);
    memcpy(out, P, nw * 8);
    free(P);
    return 0;
}

uint64_t gf2o_residue(const uint64_t *a, uint64_t abits, uint64_t q, int d) {
    size_t wa = WORDS(abits);
    uint64_t x64 = 1, b = 2, e = 64;
    while (e) {
        if (e & 1) x64 = pmulmod64(x64, b, q, d);
        b = pmulmod64(b, b, q, d);
        e >>= 1;
    }
    uint64_t r = 0;
    for (size_t w = wa; w-- > 0;) {
        uint64_t v = a[w];
        if (w == wa - 1 && (abits & 63)) v &= (~0ULL) >> (64 - (abits & 63));
        r = pmulmod64(r, x64, q, d) ^ pmod64(v, q, d);
    }
    return r;
}

uint64_t gf2o_find_irreducible(int d, uint64_t seed) {
    for (uint64_t low = seed | 1;; low += 2) {
        uint64_t q = ((uint64_t)1 << d) | (low & (((uint64_t)1 << d) - 1));
        if (gf2o_is_irreducible_small(q, d)) return q;
    }
}

/* ======================================================================
 * GCD and certificate verification
 * ==================================================================== */
int64_t gf2o_deg(const uint64_t *w, size_t nw) {
    for (size_t i = nw; i-- > 0;)
        if (w[i]) return (int64_t)(i * 64 + 63 - __builtin_clzll(w[i]));
    return -1;
}

/* A ^= B << sh with A of aw words */
static void pxor_shift(uint64_t *A, size_t aw, const uint64_t *B, int64_t bdeg, uint64_t sh) {
    size_t ws = (size_t)(sh >> 6), bw = (size_t)(bdeg >> 6) + 1;
    unsigned bs = (unsigned)(sh & 63);
    for (size_t j = bw; j-- > 0;) {
        uint64_t v = B[j];
        A[j + ws] ^= v << bs;
        if (bs && j + ws + 1 < aw) A[j + ws + 1] ^= v >> (64 - bs);
    }
}

int64_t gf2o_gcd(const uint64_t *a, size_t aw, const uint64_t *b, size_t bw,
                 uint64_t *g, size_t gcap) {
    size_t nw = (aw > bw ? aw : bw) + 1;
    uint64_t *A = (uint64_t *)calloc(nw, 8), *B = (uint64_t *)calloc(nw, 8);
    if (aw) memcpy(A, a, aw * 8);
    if (bw) memcpy(B, b, bw * 8);
    int64_t da = gf2o_deg(A, nw), db = gf2o_deg(B, nw);
    if (da < db) { uint64_t *t = A; A = B; B = t; int64_t td = da; da = db; db = td; }
    while (db >= 0) {
        while (da >= db) {
            pxor_shift(A, nw, B, db, (uint64_t)(da - db));
            /* degree can only drop: rescan from the old top word */
            size_t top = (size_t)(da >> 6) + 1;
            da = gf2o_deg(A, top);
            if (da < 0) break;
        }
        uint64_t *t = A; A = B; B = t;
        int64_t td = da; da = db; db = td;
    }
    if (g && gcap) {
        memset(g, 0, gcap * 8);
        if (da >= 0) {
            size_t gw = (size_t)(da >> 6) + 1;
            memcpy(g, A, (gw < gcap ? gw : gcap) * 8);
        }
    }
    free(A); free(B);
    return da;
}

/* r = a mod p (in place on a buffer of aw words) */
static void pmod_words(uint64_t *a, size_t aw, const uint64_t *p, int64_t dp) {
    int64_t da = gf2o_deg(a, aw);
    while (da >= dp) {
        pxor_shift(a, aw, p, dp, (uint64_t)(da - dp));
        da = gf2o_deg(a, (size_t)(da >> 6) + 1);
    }
}

/* x^e mod P (P of degree dp >= 1), result in out (words(dp) words) */
static void powx_mod(uint64_t e, const uint64_t *P, int64_t dp, uint64_t *out) {
    size_t pw = WORDS((uint64_t)dp), tw = 2 * pw + 2;
    uint64_t *acc = (uint64_t *)calloc(tw, 8), *base = (uint64_t *)calloc(tw, 8);
    uint64_t *tmp = (uint64_t *)calloc(tw + 2, 8);
    acc[0] = 1;
    base[0] = 2;
    pmod_words(acc, tw, P, dp);
    pmod_words(base, tw, P, dp);
    int top = 63;
    while (top >= 0 && !((e >> top) & 1)) top--;
    for (int bit = top; bit >= 0; bit--) {
        /* acc = acc^2 */
        memset(tmp, 0, (tw + 2) * 8);
        for (size_t i = 0; i < pw; i++) {
            tmp[2 * i] = spread32((uint32_t)acc[i]);
            tmp[2 * i + 1] = spread32((uint32_t)(acc[i] >> 32));
        }
        pmod_words(tmp, tw, P, dp);
        memcpy(acc, tmp, tw * 8);
        if ((e >> bit) & 1) {                   /* acc *= x */
            uint64_t c = 0;
            for (size_t i = 0; i < tw; i++) { uint64_t v = acc[i]; acc[i] = (v << 1) | c; c = v >> 63; }
            pmod_words(acc, tw, P, dp);
        }
    }
    memcpy(out, acc, pw * 8);
    free(acc); free(base); free(tmp);
}

int gf2o_verify_factor(uint64_t r, uint64_t s, const uint64_t *P, size_t pw) {
    int64_t d = gf2o_deg(P, pw);
    if (d < 1 || !(P[0] & 1)) return 0;
    size_t dw = WORDS((uint64_t)d);
    uint64_t *xr = (uint64_t *)calloc(dw + 1, 8), *xs = (uint64_t *)calloc(dw + 1, 8);
    powx_mod(r, P, d, xr);
    powx_mod(s, P, d, xs);
    int ok = 1;
    xr[0] ^= 1;
    for (size_t i = 0; i < dw; i++) if (xr[i] != xs[i]) ok = 0;
    free(xr); free(xs);
    if (!ok) return 0;
    /* irreducibility (Rabin): x^(2^d) == x and gcd(x^(2^(d/p)) - x, P) = 1 */
    if (d > 20000) return 1;                     /* divisibility only */
    size_t tw = 2 * dw + 2;
    uint64_t *h = (uint64_t *)calloc(tw, 8), *t = (uint64_t *)calloc(tw + 2, 8);
    uint64_t ps[16];
    int np = prime_factors((uint64_t)d, ps);
    uint64_t **snap = (uint64_t **)calloc(np, sizeof(uint64_t *));
    h[0] = 2;
    pmod_words(h, tw, P, d);
    for (int64_t k = 1; k <= d; k++) {
        memset(t, 0, (tw + 2) * 8);
        for (size_t i = 0; i < dw; i++) {
            t[2 * i] = spread32((uint32_t)h[i]);
            t[2 * i + 1] = spread32((uint32_t)(h[i] >> 32));
        }
        pmod_words(t, tw, P, d);
        memcpy(h, t, tw * 8);
        for (int i = 0; i < np; i++)
            if ((uint64_t)k == (uint64_t)d / ps[i]) {
                snap[i] = (uint64_t *)calloc(dw + 1, 8);
                memcpy(snap[i], h, dw * 8);
            }
    }
    h[0] ^= 2;
    if (gf2o_deg(h, dw) >= 0) ok = 0;
    for (int i = 0; i < np; i++) {
        if (ok && snap[i]) {
            snap[i][0] ^= 2;
            if (gf2o_gcd(snap[i], dw, P, pw, 0, 0) != 0) ok = 0;
        }
        free(snap[i]);
    }
    free(snap); free(h); free(t);
    return ok;
}

/* ======================================================================
 * Equal-degree factorization and the per-degree scan oracle
 * ==================================================================== */
/* (a * b) mod p for word arrays of pw words (p of degree dp) */
static void mulmod_words(const uint64_t *a, const uint64_t *b, const uint64_t *p, int64_t dp,
                         uint64_t *out, size_t pw) {
    size_t tw = 2 * pw + 2;
    uint64_t *t = (uint64_t *)calloc(tw, 8);
    gf2o_mul_comb(a, (uint64_t)pw * 64, b, (uint64_t)pw * 64, t);
    pmod_words(t, tw, p, dp);
    memcpy(out, t, pw * 8);
    free(t);
}

static int64_t edf_rec(const uint64_t *t, size_t tw, uint64_t d, uint64_t *best, size_t cap,
                       int64_t *best_set) {
    int64_t dt = gf2o_deg(t, tw);
    if (dt <= 0) return 0;
    size_t w = WORDS((uint64_t)dt + 1);
    if ((uint64_t)dt == d) {
        /* keep the numerically least */
        int less = !*best_set;
        if (!less) {
            for (size_t i = cap; i-- > 0;) {
                uint64_t x = i < w ? t[i] : 0;
                if (x != best[i]) { less = x < best[i]; break; }
            }
        }
        if (less) { memset(best, 0, cap * 8); memcpy(best, t, (w < cap ? w : cap) * 8); *best_set = 1; }
        return 1;
    }
    size_t pw = WORDS((uint64_t)dt);
    uint64_t *a = (uint64_t *)calloc(pw + 1, 8), *tr = (uint64_t *)calloc(pw + 1, 8);
    uint64_t *p = (uint64_t *)calloc(pw + 1, 8), *g = (uint64_t *)calloc(w + 1, 8);
    uint64_t *tt = (uint64_t *)calloc(w + 1, 8);
    memcpy(tt, t, w * 8);
    int64_t ok = 0;
    for (uint64_t i = 1; i < 4 * (uint64_t)dt + 64 && !ok; i++) {
        /* a = x^i + 1 reduced mod t */
        uint64_t *e = (uint64_t *)calloc(WORDS(i + 1) + pw + 2, 8);
        e[i >> 6] |= 1ULL << (i & 63);
        e[0] ^= 1;
        pmod_words(e, WORDS(i + 1) + pw + 2, tt, dt);
        memcpy(a, e, pw * 8);
        free(e);
        memcpy(tr, a, pw * 8);
        memcpy(p, a, pw * 8);
        for (uint64_t k = 1; k < d; k++) {
            mulmod_words(p, p, tt, dt, p, pw);
            for (size_t j = 0; j < pw; j++) tr[j] ^= p[j];
        }
        int64_t dg = gf2o_gcd(tr, pw, tt, w, g, w + 1);
        if (dg > 0 && dg < dt) {
            /* split: g and t / g */
            size_t gw = WORDS((uint64_t)dg + 1);
            uint64_t *rem = (uint64_t *)calloc(w + 2, 8), *q = (uint64_t *)calloc(w + 2, 8);
            memcpy(rem, tt, w * 8);
            int64_t dr = dt;
            while (dr >= dg) {
                uint64_t sh = (uint64_t)(dr - dg);
                q[sh >> 6] |= 1ULL << (sh & 63);
                pxor_shift(rem, w + 2, g, dg, sh);
                dr = gf2o_deg(rem, w + 2);
            }
            edf_rec(g, gw, d, best, cap, best_set);
            edf_rec(q, w + 1, d, best, cap, best_set);
            free(rem); free(q);
            ok = 1;
        }
    }
    free(a); free(tr); free(p); free(g); free(tt);
    return ok;
}

size_t gf2o_edf_least(const uint64_t *t, size_t tw, uint64_t d, uint64_t *out, size_t cap) {
    int64_t dt = gf2o_deg(t, tw);
    if (dt < (int64_t)d || d == 0) return 0;
    memset(out, 0, cap * 8);
    if (d <= 24) {
        for (uint64_t m = (1ULL << d) | 1; m < (2ULL << d); m += 2) {
            if (!gf2o_is_irreducible_small(m, (int)d)) continue;
            uint64_t mv[1] = {m};
            uint64_t g[2] = {0, 0};
            if (gf2o_gcd(mv, 1, t, tw, g, 2) == (int64_t)d) { out[0] = m; return 1; }
        }
        return 0;
    }
    int64_t set = 0;
    edf_rec(t, tw, d, out, cap, &set);
    return set ? WORDS(d + 1) : 0;
}

int gf2o_scan_one(uint64_t r, uint64_t s, uint64_t k0, uint64_t maxd, uint64_t *d,
                  uint64_t *mask, size_t cap, uint64_t *gdeg) {
    size_t nw = WORDS(r), tw = WORDS(r + 1);
    uint64_t *T = (uint64_t *)calloc(tw, 8), *F = (uint64_t *)calloc(nw, 8);
    uint64_t *Fx = (uint64_t *)calloc(nw, 8), *g = (uint64_t *)calloc(tw + 1, 8);
    T[0] |= 1; T[s >> 6] |= 1ULL << (s & 63); T[r >> 6] |= 1ULL << (r & 63);
    F[0] = 2;
    int kind = 2;
    for (uint64_t k = 1; k <= maxd; k++) {
        gf2o_square_fold_steps(F, r, s, 1);
        memcpy(Fx, F, nw * 8);
        Fx[0] ^= 2;
        int64_t gd = gf2o_gcd(Fx, nw, T, tw, g, tw + 1);
        if (gd > 0) {
            if (k <= k0) { kind = 0; break; }
            *d = k;
            if (gdeg) *gdeg = (uint64_t)gd;
            if ((uint64_t)gd == k) {
                memset(mask, 0, cap * 8);
                size_t gw = WORDS((uint64_t)gd + 1);
                memcpy(mask, g, (gw < cap ? gw : cap) * 8);
            } else {
                gf2o_edf_least(g, WORDS((uint64_t)gd + 1), k, mask, cap);
            }
            kind = 1;
            break;
        }
    }
    free(T); free(F); free(Fx); free(g);
    return kind;
}
