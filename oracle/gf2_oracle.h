/*
 * gf2_oracle.h -- CPU restatement of the reference's algorithms for the
 * GF(2) trinomial smallest-factor path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * Conventions (same as the reference): polynomials over GF(2) are
 * little-endian uint64_t word arrays, bit i of word w is the coefficient
 * of x^(64w+i); residues mod T = x^r + x^s + 1 have nw = ceil(r/64) words
 * with all bits >= r zero.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   field / sieve      brute-force irreducible enumeration (independent
 *                      code path), irreducible counts by the Moebius
 *                      formula, the reference's sieve_ref output (md5 of
 *                      r=132049 d<=20) and published counts.
 *   square_fold        naive schoolbook square + bit-serial reduction,
 *                      F_31 == x at r=31 s=3, md5 of F_32 at r=136279841
 *                      s=474 produced by the reference stage-2 driver,
 *                      linearity.
 *   gf64 / Cantor FFT  bit-serial field multiply, direct evaluation of the
 *                      chunk polynomial at the Cantor points (small m),
 *                      schoolbook products, the reference host engine
 *                      built from its own header (oracle/_ref).
 *   gcd / verify       brute force on tiny inputs, planted common factors.
 */
#ifndef GF2_ORACLE_H
#define GF2_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- GF(2^d), d <= 32 (reference: src/gf2_small_field.h:30-135) ---- */
uint64_t gf2o_clmul32(uint32_t a, uint32_t b);
uint32_t gf2o_fmul(uint32_t a, uint32_t b, uint64_t q, int d);
int      gf2o_is_irreducible(uint64_t q, int d);             /* d <= 32 */
uint64_t gf2o_find_primitive_poly(int d);                    /* smallest */
uint64_t gf2o_count_irreducible(int d);                      /* Moebius */

/* ---- coarse sieve (reference: src/sieve_ref.cpp:58-160) ----
 * best[0..smax] is min-updated with (d<<48)|mask; caller initialises.
 * classes[d] (if non-null, >= dmax+1 entries) receives the number of
 * surviving cyclotomic classes per degree.  nthreads >= 1. */
int gf2o_sieve(uint64_t r, int dmin, int dmax, uint64_t smax,
               uint64_t *best, uint64_t *classes, int nthreads);
int gf2o_brute_force(uint64_t r, int depth, uint64_t smax, uint64_t *best);
/* 1 iff the degree-d polynomial `mask` (d <= 62) divides x^r + x^s + 1 */
int gf2o_divides_trinomial_small(uint64_t r, uint64_t s, uint64_t mask, int d);
int gf2o_is_irreducible_small(uint64_t mask, int d);         /* d <= 62 */

/* ---- residues mod x^r + x^s + 1 (reference: src/gf2_trinomial_cuda.h
 *      :22-74, :182-198; src/tsfactor.cu:128-186) ---- */
void gf2o_square_fold(const uint64_t *in, uint64_t *out, uint64_t r,
                      uint64_t s);
void gf2o_square_fold_steps(uint64_t *res, uint64_t r, uint64_t s,
                            uint64_t steps);
/* word-level two-pass fold of v (bits < `bits`, bits <= 2r-1) in place;
 * v must hold words(max(bits, r)) words */
void gf2o_reduce_fold(uint64_t *v, uint64_t bits, uint64_t r, uint64_t s);
void gf2o_reduce_bitserial(uint64_t *v, uint64_t bits, uint64_t r,
                           uint64_t s);
void gf2o_mulx_mod(const uint64_t *in, uint64_t *out, uint64_t r,
                   uint64_t s);

/* ---- GF(2)[x] products; out has words(abits+bbits)+1 words ---- */
size_t gf2o_mul_out_words(uint64_t abits, uint64_t bbits);
void gf2o_mul_naive(const uint64_t *a, uint64_t abits, const uint64_t *b,
                    uint64_t bbits, uint64_t *out);
void gf2o_mul_comb(const uint64_t *a, uint64_t abits, const uint64_t *b,
                   uint64_t bbits, uint64_t *out);

/* ---- GF(2^64) and the Cantor additive FFT (reference:
 *      src/gf2_cantor_core.h:30-110, src/gf2_cantor_engine.h:52-242) ---- */
uint64_t gf2o_gf64_mul_bitserial(uint64_t a, uint64_t b);
uint64_t gf2o_gf64_mul(uint64_t a, uint64_t b);
int  gf2o_cantor_basis(uint64_t b[64]);                      /* -> len */
void gf2o_cantor_point(const uint64_t *basis, uint64_t idx, uint64_t *pt);
void gf2o_fft_fwd(uint64_t *f, int m, const uint64_t *basis);
void gf2o_fft_inv(uint64_t *f, int m, const uint64_t *basis);
/* evaluate sum_k f[k] X^k (n coefficients) at X = pt via Horner */
uint64_t gf2o_poly_eval64(const uint64_t *f, size_t n, uint64_t pt);
int  gf2o_fft_mul(const uint64_t *a, uint64_t abits, const uint64_t *b,
                  uint64_t bbits, uint64_t *out);
/* full product then fold mod x^r + x^s + 1 into out (nw words) */
int  gf2o_mulmod(const uint64_t *a, uint64_t abits, const uint64_t *b,
                 uint64_t bbits, uint64_t r, uint64_t s, uint64_t *out);

/* value of the bit array a (abits bits) modulo q (degree d <= 62), by
 * Horner over 64-bit words (reference: ModP::residue,
 * src/gf2_cantor_engine.h:316-324) */
uint64_t gf2o_residue(const uint64_t *a, uint64_t abits, uint64_t q, int d);
/* smallest irreducible of degree d with low bits >= seed|1 (reference:
 * find_irreducible_modp, src/gf2_cantor_engine.h:327-334) */
uint64_t gf2o_find_irreducible(int d, uint64_t seed);

/* ---- GCD and verification (reference: src/gf2_gcd.h:48-145) ---- */
int64_t gf2o_deg(const uint64_t *w, size_t nw);
/* naive Euclid; returns deg(gcd) (-1 if both zero); writes
 * min(gcap, words) words of the monic gcd into g (g may be NULL) */
int64_t gf2o_gcd(const uint64_t *a, size_t aw, const uint64_t *b,
                 size_t bw, uint64_t *g, size_t gcap);
/* 1 iff P (degree d, pw words) divides x^r + x^s + 1 and is irreducible */
int gf2o_verify_factor(uint64_t r, uint64_t s, const uint64_t *P,
                       size_t pw);

/* numerically least irreducible factor of t whose factors all have
 * degree d: brute-force enumeration of degree-d masks for d <= 24 (the
 * reference's edf_least_enum, src/gf2_gcd.h:137-145), trace splitting
 * with the elements x^i + 1 above that.  Returns words written (0 if
 * none). */
size_t gf2o_edf_least(const uint64_t *t, size_t tw, uint64_t d, uint64_t *out, size_t cap);

/* Per-degree deep-scan oracle (reference: oracle_scan,
 * src/trinomial_scan.cu:393-421): F_k = x^(2^k) mod T, first k with
 * gcd(F_k + x, T) != 1.  kind: 0 not a survivor (hit at k <= k0),
 * 1 found (d, least mask in mask[]), 2 none up to maxd.  Returns kind. */
int gf2o_scan_one(uint64_t r, uint64_t s, uint64_t k0, uint64_t maxd, uint64_t *d,
                  uint64_t *mask, size_t cap, uint64_t *gdeg);

#ifdef __cplusplus
}
#endif
#endif
