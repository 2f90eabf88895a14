/*
 * gf2t.h -- C ABI of the MI355X (gfx950) GF(2) trinomial smallest-factor
 * engine: Zech-log coarse sieve, squaring+fold Frobenius chains and the
 * Cantor additive-FFT multiply, on one GPU.
 *
 * Conventions
 *  - Polynomials over GF(2) are little-endian uint64_t word arrays: bit i
 *    of word w is the coefficient of x^(64w+i).  A residue modulo
 *    T = x^r + x^s + 1 has nw = ceil(r/64) words and all bits >= r zero,
 *    on input and on output.  Valid trinomials: r >= 3, 1 <= s, 2s < r.
 *  - Pointers named d* are device pointers owned by the caller; h* are
 *    host pointers.  Contexts own their scratch and plans.
 *  - Every call returns 0 on success and a negative GF2T_E* code on
 *    failure, never exit(); gf2t_last_error() describes the last failure
 *    of the calling thread.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *    Calls are asynchronous unless stated; results are ordered on the
 *    stream.  One host thread per context.
 *
 * Reference citations are file:line in nclvrps/mersenne_trinomial.
 */
#ifndef GF2T_H
#define GF2T_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GF2T_OK 0
#define GF2T_EINVAL (-1)   /* bad argument (reference: exit(1) paths) */
#define GF2T_EHIP (-2)     /* HIP runtime error */
#define GF2T_ENOMEM (-3)   /* device allocation failed */
#define GF2T_ERANGE (-4)   /* size beyond a supported limit */
#define GF2T_EINTERNAL (-5)/* internal consistency failure */

/* ------------------------------------------------------------------ *
 * Runtime, memory, streams, events.  The reference uses the CUDA
 * runtime directly (CUCHK, src/gf2_cantor_cuda.h:65-69); these are the
 * plain-pointer equivalents so a host without HIP headers can drive the
 * library.
 * ------------------------------------------------------------------ */
const char *gf2t_last_error(void);
int gf2t_version(void);
int gf2t_init(int device);
int gf2t_device_count(int *count);
int gf2t_device_name(char *buf, size_t cap);
int gf2t_malloc(void **dptr, size_t bytes);
int gf2t_free(void *dptr);
int gf2t_mem_info(size_t *free_bytes, size_t *total_bytes);
int gf2t_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int gf2t_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int gf2t_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int gf2t_memset(void *dptr, int value, size_t bytes, void *stream);
int gf2t_stream_create(void **stream);
int gf2t_stream_destroy(void *stream);
int gf2t_stream_sync(void *stream);
int gf2t_device_sync(void);
int gf2t_event_create(void **event);
int gf2t_event_destroy(void *event);
int gf2t_event_record(void *event, void *stream);
int gf2t_event_elapsed_ms(void *start, void *stop, float *ms);
/* HIP graphs of call sequences.  gf2t_graph_begin starts capturing what
 * the calling thread issues on `stream` (a stream from gf2t_stream_create,
 * not NULL); gf2t_graph_end stops and instantiates it.  Each
 * gf2t_graph_launch replays the same launches, with the same pointers and
 * sizes, on any stream.  Only kernel-only calls may be captured: the
 * residue-ring, transform, product and slab entry points (gf2t_square_*,
 * gf2t_fold, gf2t_mulx_xor*, gf2t_set_*, gf2t_xor_*, gf2t_copy_slab,
 * gf2t_horner_slab, gf2t_fft_fwd / _inv, gf2t_mul, gf2t_mulmod,
 * gf2t_mulmod_slab, gf2t_mulmod_noalloc); not the sieve, the host-buffer
 * services, copies to or from the host, gf2t_memset or
 * gf2t_ddf_init_slab.  The reference launches every kernel from the host
 * (src/tsfactor.cu:278-336); the deep scan replays one graph per block of
 * m degrees.  gf2t_graph_nodes reports the captured node count. */
typedef struct gf2t_graph gf2t_graph;
int gf2t_graph_begin(void *stream);
int gf2t_graph_end(void *stream, gf2t_graph **graph);
int gf2t_graph_launch(gf2t_graph *graph, void *stream);
int gf2t_graph_nodes(const gf2t_graph *graph, size_t *nodes);
void gf2t_graph_destroy(gf2t_graph *graph);

/* 16-byte vector copy kernel (dst, src 16-byte aligned, bytes % 16 == 0):
 * the stream-copy reference the HBM roofline figures are calibrated
 * against; no reference counterpart. */
int gf2t_stream_copy(void *dst, const void *src, size_t bytes, void *stream);

/* ------------------------------------------------------------------ *
 * (b) Residue ring GF(2)[x]/(x^r + x^s + 1)
 * ------------------------------------------------------------------ */

/* dout = din^2 mod T in one pass (src != dst).  Replaces k_square_fused
 * (src/tsfactor.cu:173-186) as used by TsGPU::square_slot
 * (src/tsfactor.cu:278-300) and TrinGPU::square_A
 * (src/gf2_trinomial_cuda.h:149-154). */
int gf2t_square_fold(const uint64_t *din, uint64_t *dout, uint64_t r,
                     uint64_t s, void *stream);

/* dres = dres^(2^steps) mod T; dtmp is nw words of scratch.  The
 * per-s Frobenius chain of trinomial_stage2 run (k_set_x then k x
 * square_A, src/trinomial_stage2.cu:228-235). */
int gf2t_square_fold_steps(uint64_t *dres, uint64_t *dtmp, uint64_t r,
                           uint64_t s, uint64_t steps, void *stream);

/* Batched Frobenius chains: residue i (r[i], s[i], nw_i words in
 * dres[i], scratch dtmp[i]) is squared steps[i] times; steps may differ
 * (uneven frontiers) and r may differ between residues.  Results end in
 * dres[i] (buffers 16-byte aligned, as gf2t_malloc returns them).  All
 * arrays are host arrays of n entries.  No reference
 * counterpart: it batches the per-s loop of src/trinomial_scan.cu:364-
 * 366 / src/tsfactor.cu:294-300 across s values. */
int gf2t_square_fold_batch(uint64_t *const *dres, uint64_t *const *dtmp,
                           const uint64_t *r, const uint64_t *s,
                           const uint64_t *steps, int n, void *stream);

/* Tuning knobs of the kernels (the autotuner's handle; the reference
 * reads GF2C_* environment variables, src/gf2_cantor_cuda.h:24-61):
 *   "sq_fuse"  1 = two squarings per launch in gf2t_square_fold_steps and
 *              gf2t_square_fold_batch while every live chain has two
 *              steps left; 0 = one squaring per launch
 *   "sq_ppt"   output pairs per lane in the one-step batched squaring
 *              (1, 2, 4)
 *   "sq_order" 1 = XCD-aware tile order in the one-step batched squaring,
 *              0 = linear
 *   "sq_nt"    1 = non-temporal (streaming) stores in the batched squaring
 *              and in gf2t_stream_copy
 *   "gcd_dev_bits" products inside the host gcds (gf2t_gcd_host, the
 *              deep scan) go to the device when both operands have at
 *              least this many bits; 0 = host Karatsuba only
 *   "gcd_base_bits" the host half-gcd runs its base case (rounds on the
 *              top 128 coefficients) below this size in bits (128..65536;
 *              default 3072)
 *   "gcd_school_words" the host Karatsuba's schoolbook base in words
 *              (8..512; default 48)
 *   "fft_ept"  elements of a transform tile per thread (2, 4, 8, 16;
 *              default 2): fewer threads per tile
 *   "fft_low_bits" low index bits kept local in every transform pass,
 *              i.e. contiguous elements per HBM run of a tile (4, 5 or 6;
 *              default 4): longer runs, but fewer window bits per pass
 *              and so more passes.  Plans made before keep theirs
 *   "hm_zc_words" products of the host-buffer service (gf2t_mat2n_host,
 *              gf2t_mat2_host) whose operands and results take at most
 *              this many words together are read and written by the GPU
 *              in pinned host memory directly, with no copy each way
 *              (default 65536; 0 = always copy)
 *   "scan_stagger" gf2t_scan_run splits each cohort into this many parts
 *              that start their device work one after another, so the
 *              gcds of one part overlap the device work of the next
 *              (not with a progress callback); 0 = up to 4 parts from
 *              r = 2^22 on, -1 = never
 * Environment defaults: GF2T_SQ_FUSE (1), GF2T_SQ_PPT (2), GF2T_SQ_ORDER (0),
 * GF2T_SQ_NT (1), GF2T_GCD_DEV_BITS (2^16, set by gf2t_init; 0 before it).
 * GF2T_NO_VPCLMUL=1 keeps the host gcds' schoolbook products on 128-bit
 * PCLMULQDQ where the CPU also has AVX-512 VPCLMULQDQ (read once). */
int gf2t_tune(const char *key, int64_t value);
/* The value a knob has now (the default until set): the keys above, and
 * "fft_tile_<m>" -- the tile of the 2^m-point transforms' passes is 2^T
 * elements, T in 6..13 (set with gf2t_tune("fft_tile_<m>", T), 0 = the
 * default: 13 from m = 19 up, 11 for m = 14..18, m below 14).  A smaller
 * tile spreads a small transform over more workgroups at the cost of more
 * passes; the reference's split knobs are gf2c_taylor_lv / gf2c_bfly_lv
 * (src/gf2_cantor_cuda.h:24-61).  The autotuner (mersenne_trinomial_amd/
 * autotune.py) measures these and installs its choices. */
int gf2t_tune_get(const char *key, int64_t *value);

/* Slab form: residues i = 0..n-1 at din + i*pitch (pitch >= nw words),
 * each squared once mod x^r + x^s_i + 1 into dout + i*pitch; ds is a
 * DEVICE array of the n values s_i (each 1 <= s_i, 2 s_i < r).  The
 * batched per-degree step of the deep scan (square_A over many s,
 * src/trinomial_scan.cu:364-366). */
int gf2t_square_fold_slab(const uint64_t *din, uint64_t *dout, size_t pitch,
                          uint64_t r, const uint64_t *ds, int n, void *stream);

/* Fold a product of src_bits <= 2r-1 bits (dsrc, words(src_bits) words)
 * into dout (nw words) modulo T.  Replaces TrinGPU::reduce / TsGPU::
 * reduce_into (two k_fold_pass launches, src/gf2_trinomial_cuda.h:63-74,
 * :136-146; src/tsfactor.cu:323-331). */
int gf2t_fold(const uint64_t *dsrc, uint64_t src_bits, uint64_t *dout,
              uint64_t r, uint64_t s, void *stream);

/* dst ^= x*src mod T (k_mulx_xor, src/tsfactor.cu:109-126). */
int gf2t_mulx_xor(uint64_t *ddst, const uint64_t *dsrc, uint64_t r,
                  uint64_t s, void *stream);

/* p[0..nw) = x (k_set_x), = 1 (k_set_one), p ^= x (k_xor_x)
 * (src/gf2_trinomial_cuda.h:76-84, src/tsfactor.cu:198-202). */
int gf2t_set_x(uint64_t *dp, size_t nw, void *stream);
int gf2t_set_one(uint64_t *dp, size_t nw, void *stream);
int gf2t_xor_x(uint64_t *dp, void *stream);

/* ------------------------------------------------------------------ *
 * (c) Cantor additive FFT over GF(2^64) and GF(2)[x] products
 * ------------------------------------------------------------------ */

/* Cantor basis b[0]=1, b[i]^2+b[i]=b[i-1]; returns the chain length
 * (64) in *len (CantorBasis::build/verify, src/gf2_cantor_engine.h:74-
 * 121). */
int gf2t_cantor_basis(uint64_t *hbasis, int *len);

typedef struct gf2t_fft_plan gf2t_fft_plan;
/* Plan for 2^m points, 1 <= m <= 27 (CudaFFT::init, src/gf2_cantor_cuda.
 * h:365-375; the plan computes twiddles on the fly, no table). */
int gf2t_fft_plan_create(int m, gf2t_fft_plan **plan);
int gf2t_fft_plan_destroy(gf2t_fft_plan *plan);
/* number of HBM passes of one transform, for reporting */
int gf2t_fft_plan_passes(const gf2t_fft_plan *plan, int *fwd_passes,
                         int *inv_passes);
/* In-place forward / inverse transform of 2^m elements at df; output is
 * bit-identical to CantorFFT::fwd/inv (src/gf2_cantor_engine.h:146-187)
 * and CudaFFT::fwd/inv (src/gf2_cantor_cuda.h:497-520). */
int gf2t_fft_fwd(const gf2t_fft_plan *plan, uint64_t *df, void *stream);
int gf2t_fft_inv(const gf2t_fft_plan *plan, uint64_t *df, void *stream);

/* Number of output words of a product: words(abits+bbits)+1
 * (mul_out_words, src/gf2_cantor_engine.h:201-203). */
size_t gf2t_mul_out_words(uint64_t abits, uint64_t bbits);

/* Workspace (bytes) that gf2t_mul / gf2t_mulmod need for these sizes. */
size_t gf2t_mul_workspace_bytes(uint64_t abits, uint64_t bbits);

/* dout = a*b over GF(2)[x] (gf2t_mul_out_words words, fully written);
 * dwork is caller-owned scratch of gf2t_mul_workspace_bytes bytes.
 * Replaces gf2x_fft_mul_buf (src/gf2_cantor_engine.h:207-229) and
 * TrinGPU::mul_dev / TsGPU::mul_full (src/gf2_trinomial_cuda.h:157-169,
 * src/tsfactor.cu:309-321). */
int gf2t_mul(const uint64_t *da, uint64_t abits, const uint64_t *db,
             uint64_t bbits, uint64_t *dout, void *dwork, void *stream);

/* dout (nw words) = a*b mod x^r + x^s + 1 with abits + bbits - 1 <=
 * 2r - 1.  Replaces TrinGPU::accumulate's mul_dev+reduce and
 * TsGPU::mul_A_by_C (src/gf2_trinomial_cuda.h:172-177,
 * src/tsfactor.cu:333-336).  dout may alias da or db. */
int gf2t_mulmod(const uint64_t *da, uint64_t abits, const uint64_t *db,
                uint64_t bbits, uint64_t r, uint64_t s, uint64_t *dout,
                void *dwork, void *stream);

/* dout (nw words) = a*b mod x^r + x^s + 1, as gf2t_mulmod, but with no
 * workspace argument and no allocation: the transforms and the product
 * live in a library-owned device array of 80 MiB per GPU
 * (2 x 2^22 transform points + a 2^27-bit product), reserved when the
 * library is loaded and kept until the process ends.  Limit: the two
 * operands together hold at most 2^22 + 1 32-bit chunks (abits + bbits
 * up to about 2^27, e.g. r up to 2^26 for two r-bit operands); larger
 * products return GF2T_ERANGE.  Calls that overlap on one GPU share the
 * array and race; calls ordered on one stream, or synchronized, do not.
 * dout may alias da or db.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that must not call hipMalloc). */
int gf2t_mulmod_noalloc(const uint64_t *da, uint64_t abits, const uint64_t *db,
                        uint64_t bbits, uint64_t r, uint64_t s, uint64_t *dout,
                        void *stream);

/* n products of different sizes, each folded: dout[j] (nw words) =
 * a_j * b_j mod x^r + x^s[j] + 1 (da, db, dout: host arrays of device
 * pointers; abits, bbits, s: host arrays), with dwork of
 * gf2t_mulmod_batch_workspace_bytes(n, abits, bbits) bytes.  Every pair
 * keeps its own transform size, and the batch runs as one set of
 * launches: the i-th pass of every pair's forward (inverse) transforms is
 * one grouped launch, and the pointwise products, overlap-adds and folds
 * are one launch each.  Synchronous with respect to the host only for
 * its small descriptor upload.  No reference counterpart: the reference
 * multiplies one pair at a time (gf2x_fft_mul_buf,
 * src/gf2_cantor_engine.h:207-229). */
size_t gf2t_mulmod_batch_workspace_bytes(int n, const uint64_t *abits, const uint64_t *bbits);
int gf2t_mulmod_batch(int n, const uint64_t *const *da, const uint64_t *abits,
                      const uint64_t *const *db, const uint64_t *bbits, uint64_t r,
                      const uint64_t *s, uint64_t *const *dout, void *dwork, void *stream);

/* Slab form of the interval accumulation: for i = 0..n-1,
 * out_i = a_i * (b_i + [b_add_x] x) mod x^r + x^s_i + 1, with a_i at
 * da + i*apitch, b_i at db + i*bpitch, out_i at dout + i*opitch (r-bit
 * operands; dout may equal da), ds a DEVICE array of s_i, dwork of
 * gf2t_mulmod_slab_workspace_bytes(r, n) bytes.  The batched
 * TrinGPU::accumulate (ACC <- ACC * (F + x), src/gf2_trinomial_cuda.h:
 * 172-177) over many s in one set of launches. */
size_t gf2t_mulmod_slab_workspace_bytes(uint64_t r, int n);
int gf2t_mulmod_slab(const uint64_t *da, size_t apitch, const uint64_t *db,
                     size_t bpitch, int b_add_x, uint64_t r, const uint64_t *ds,
                     uint64_t *dout, size_t opitch, int n, void *dwork,
                     void *stream);

/* ------------------------------------------------------------------ *
 * Blocked distinct-degree factorization on slabs (tsfactor's h-table:
 * TsGPU::init_h_at / square_all_step / horner_C, src/tsfactor.cu:
 * 71-126, :278-300, :302-307, :345-365), for a cohort of s values.
 * h-table of n residues: slot j (j < m) of residue i at
 * dh + (i*m + j)*slot_pitch; slot j holds e_(j+1)(x^(2^(k+1)), ...,
 * x^(2^(k+m))) for the block of degrees k+1..k+m.  ds: DEVICE array of
 * the n values s_i.  Slab arrays: residue i at base + i*pitch.
 * ------------------------------------------------------------------ */
/* Pascal build of the h-table for block start kstart (degrees kstart..
 * kstart+m-1), into dh; dh_alt is an equal-sized scratch table.
 * 1 <= m <= 64. */
int gf2t_ddf_init_slab(uint64_t *dh, uint64_t *dh_alt, size_t slot_pitch, int m,
                       uint64_t r, const uint64_t *ds, uint64_t kstart, int n,
                       void *stream);
/* n residues (pitch words apart) squared once; residue y uses
 * ds[y / group] -- with group = m this advances a whole h-table by one
 * degree (square_all_step).  No limit on n. */
int gf2t_square_fold_slab_g(const uint64_t *din, uint64_t *dout, size_t pitch,
                            uint64_t r, const uint64_t *ds, size_t n, int group,
                            void *stream);
/* The same advanced by two squarings in one pass (dout = din^4; din is
 * not changed; pitch even and both slabs 16-byte aligned).  No reference
 * counterpart: two of square_all_step (src/tsfactor.cu:278-300) fused,
 * for the m squarings of a block. */
int gf2t_square2_fold_slab_g(const uint64_t *din, uint64_t *dout, size_t pitch,
                             uint64_t r, const uint64_t *ds, size_t n, int group,
                             void *stream);
/* The same two squarings in work units (tiles one period apart that pass
 * window copies on, as gf2t_square_fold_batch's two-squaring launches),
 * with each group's tile layout computed on the host once: a plan holds
 * the layouts of up to ngroups groups (s values hs[g], host memory) at one
 * r; _set replaces them (at most the created count, e.g. after a cohort
 * is compacted; it waits for the stream's earlier work); destroy it only
 * when no launch that uses it is still queued.  Same result
 * as gf2t_square2_fold_slab_g with ds = hs; n / group rounded up must
 * not exceed the groups set.  No reference counterpart (see above). */
typedef struct gf2t_sq2_slab_plan gf2t_sq2_slab_plan;
int gf2t_sq2_slab_plan_create(uint64_t r, const uint64_t *hs, size_t ngroups,
                              gf2t_sq2_slab_plan **plan);
int gf2t_sq2_slab_plan_set(gf2t_sq2_slab_plan *plan, const uint64_t *hs, size_t ngroups,
                           void *stream);
int gf2t_sq2_slab_plan_destroy(gf2t_sq2_slab_plan *plan);
int gf2t_square2_fold_slab_p(const gf2t_sq2_slab_plan *plan, const uint64_t *din,
                             uint64_t *dout, size_t pitch, size_t n, int group,
                             void *stream);
/* dc_i = x^m + sum_j h_(i,j) x^(m-1-j) mod T_i = prod_t (x + x^(2^(k+1+t)))
 * (horner_C: k_horner_fused + k_horner_tail); 2m < r. */
int gf2t_horner_slab(const uint64_t *dh, size_t slot_pitch, int m, uint64_t r,
                     const uint64_t *ds, uint64_t *dc, size_t cpitch, int n,
                     void *stream);
/* slab helpers: dst_i ^= x src_i mod T_i (k_mulx_xor); p_i[0] ^= v;
 * p_i = 1; dst_i = src_i */
int gf2t_mulx_xor_slab(uint64_t *ddst, size_t dpitch, const uint64_t *dsrc,
                       size_t spitch, uint64_t r, const uint64_t *ds, int n,
                       void *stream);
int gf2t_xor_word0_slab(uint64_t *dp, size_t pitch, uint64_t v, int n, void *stream);
int gf2t_set_one_slab(uint64_t *dp, size_t pitch, size_t nw, int n, void *stream);
int gf2t_copy_slab(uint64_t *ddst, size_t dpitch, const uint64_t *dsrc,
                   size_t spitch, size_t nw, int n, void *stream);

/* Host-buffer product service: the device side of the reference's
 * GpuMulHook (TsGpuMul::mul / mat2_apply, src/tsfactor.cu:467-607;
 * interface src/gf2_gcd.h:155-172).  Host pointers in and out; the
 * library keeps one mutex-guarded stream and device buffers (grown on
 * demand, kept until exit), so any host thread may call.  Synchronous.
 *   gf2t_mul_host:  hout (gf2t_mul_out_words(abits, bbits) words) = a*b.
 *   gf2t_mat2_host: (oa, ob) = (m0 a + m1 b, m2 a + m3 b) with one set of
 *   six forward and two inverse transforms; outw >= gf2t_mul_out_words(
 *   max mbits, max(abits, bbits)), both outputs fully written. */
int gf2t_mul_host(const uint64_t *ha, uint64_t abits, const uint64_t *hb,
                  uint64_t bbits, uint64_t *hout);
int gf2t_mat2_host(const uint64_t *const hm[4], const uint64_t mbits[4],
                   const uint64_t *ha, uint64_t abits, const uint64_t *hb,
                   uint64_t bbits, uint64_t *hoa, uint64_t *hob, size_t outw);
/* The same matrix applied to npairs (1..4) vector pairs with one set of
 * four matrix transforms: pair i is (hv[2i], hv[2i+1]) of vbits[2i],
 * vbits[2i+1] bits, and its results go to hout[2i] = m0 a + m1 b and
 * hout[2i+1] = m2 a + m3 b, each outw >= gf2t_mul_out_words(max mbits,
 * max vbits) words, fully written.  A half-gcd node applies its second
 * matrix to the remainder pair and to both columns of the first matrix
 * this way.  No reference counterpart (it batches mat2_apply,
 * src/tsfactor.cu:535-607). */
int gf2t_mat2n_host(const uint64_t *const hm[4], const uint64_t mbits[4], int npairs,
                    const uint64_t *const *hv, const uint64_t *vbits, uint64_t *const *hout,
                    size_t outw);

/* ------------------------------------------------------------------ *
 * (a) Zech-log coarse sieve
 * ------------------------------------------------------------------ */
typedef struct gf2t_sieve_stats {
    int d;                    /* degree */
    uint64_t q;               /* primitive modulus used */
    uint64_t classes;         /* cyclotomic classes (= irreducibles of degree d) */
    uint64_t marking;         /* classes whose progression reaches s <= smax */
    uint64_t heavy;           /* progressions deferred to the heavy pass */
    float ms_table, ms_classes, ms_heavy;
} gf2t_sieve_stats;

/* Mark every s in [1, smax] divisible by an irreducible of degree d in
 * [dmin, dmax] (2 <= dmin <= dmax <= 36) into dbest (smax+1 uint64 on
 * the device, atomic min of (d<<48)|mask; initialise to ~0 or a prior
 * image).  Replaces run_gpu/sieve_degree_gpu (src/coarse_sieve.cu:852-
 * 918, :1119-1225).  stats (host, dmax+1 entries, may be NULL) gets one
 * record per degree.  Degrees 33..36 (coarse_sieve_wide, src/
 * gf2_wide_field.h) use 64-bit field elements and tables of about 205 GB
 * at d = 34; degrees 35 and 36 build and use the log table in 2 and 4
 * value slices of that size (the role of the reference's low-memory
 * buckets, src/coarse_sieve.cu:962-1117); the tables of degrees <= 32 are
 * freed first.  Test knobs: gf2t_tune "sieve_wide_from" (25..33) runs
 * degrees from there on the 64-bit path, "sieve_slices" (1..16) forces at
 * least that many slices.  Synchronous. */
int gf2t_sieve(uint64_t r, int dmin, int dmax, uint64_t smax,
               uint64_t *dbest, gf2t_sieve_stats *stats);

/* The sieve keeps its device tables (up to ~83 GB at dmax = 32) between
 * calls; this frees them.  No reference counterpart. */
int gf2t_sieve_release(void);

/* Smallest primitive polynomial of degree d (find_primitive_poly,
 * src/gf2_field_setup.h:49-57); 0 if d is out of range. */
uint64_t gf2t_primitive_poly(int d);

/* ------------------------------------------------------------------ *
 * Deep scan of sieve survivors (squaring chain + interval products on
 * the device, interval gcds and factor resolution on host threads)
 * ------------------------------------------------------------------ */
#define GF2T_SCAN_FOUND 1      /* "s d p<hex>"                              */
#define GF2T_SCAN_GAVE_UP 2    /* "s u": maxd reached below the Swan bound  */
#define GF2T_SCAN_PRIMITIVE 3  /* "s primitive": no factor up to the bound  */
#define GF2T_SCAN_RANGE 4      /* "s rA-B": factor in [A,B], unresolved     */
#define GF2T_SCAN_STOPPED 5    /* the progress callback stopped the scan
                                  before this s finished (no line)          */

typedef struct gf2t_scan_params {
    uint64_t k0;      /* depth the coarse sieve reached (no factor <= k0) */
    uint64_t maxd;    /* 0 = Swan bound (r/3 or r/2), else min(maxd, it) */
    uint64_t L0;      /* first interval length (0 = 64)                  */
    uint64_t Lmax;    /* interval length cap (0 = 4096)                  */
    int gcd_threads;  /* host threads for interval gcds (0 = 8)          */
    int batch;        /* max s values per device cohort (0 = by memory)  */
    int verbose;
    int block_m;      /* degrees per product: 1 = one product per degree
                         (trinomial_scan), m > 1 = blocked h-table scheme
                         (tsfactor), 0 = chosen from r                  */
    int past_bound;   /* 1 = report a factor found in the last whole block
                         of m degrees even when its degree exceeds the
                         bound, as tsfactor does (its intervals always end
                         on block boundaries, src/tsfactor.cu:1266-1273);
                         0 = report only degrees <= the bound           */
    /* Called after every interval of a cohort with the degree k_done
     * that each of the n listed s values is now known to be clear
     * through (no factor of degree <= k_done), the s values still being
     * scanned.  A scan restarted for them with k0 = k_done continues
     * exactly where this one is (the checkpoint of the reference's
     * tsfactor, src/tsfactor.cu:717-828, reduced to what resuming
     * needs).  Returning nonzero stops the scan: every s not finished
     * then gets GF2T_SCAN_STOPPED.  May be null. */
    int (*progress)(void *user, uint64_t k_done, const uint64_t *svals, size_t n);
    void *progress_user;
} gf2t_scan_params;

typedef struct gf2t_scan_result {
    uint64_t s;
    int kind;             /* GF2T_SCAN_* */
    uint64_t d;           /* least factor degree (FOUND)                 */
    uint64_t rlo, rhi;    /* RANGE bounds                                */
    uint64_t gcd_degree;  /* degree of the interval gcd that hit         */
    size_t hex_len;       /* length of the mask hex (FOUND)              */
} gf2t_scan_result;

typedef struct gf2t_scan_results gf2t_scan_results;

/* Scan every s of svals (1 <= s, 2s < r, r < 2^31).  Replaces
 * trinomial_scan's per-s loop (scan_one_s, src/trinomial_scan.cu:240-381)
 * and tsfactor's (src/tsfactor.cu:1089-1348) with cohorts of s values
 * advanced together on the device; results follow tsfactor's kinds
 * (src/tsfactor.cu:2480-2509).  Synchronous. */
int gf2t_scan_run(uint64_t r, const uint64_t *svals, size_t n,
                  const gf2t_scan_params *params, gf2t_scan_results **out);
size_t gf2t_scan_results_count(const gf2t_scan_results *res);
int gf2t_scan_results_get(const gf2t_scan_results *res, size_t i,
                          gf2t_scan_result *row, char *hexbuf, size_t cap);
/* Wall seconds the run spent waiting for the device, in interval gcds
 * and in resolving hits (summed over its cohorts; the device work of the
 * next interval overlaps the gcds), and the number of interval gcds.
 * Any pointer may be NULL.  The reference prints the same phases per s
 * (src/tsfactor.cu:1334-1346). */
int gf2t_scan_results_times(const gf2t_scan_results *res, double *dev_s, double *gcd_s,
                            double *resolve_s, uint64_t *n_gcd);
/* the certificate line of result i ("s d p<hex>", "s u", ...) */
int gf2t_scan_line(const gf2t_scan_results *res, size_t i, char *buf,
                   size_t cap);
void gf2t_scan_results_free(gf2t_scan_results *res);

/* ------------------------------------------------------------------ *
 * Deep-scan server with continuous batching.  No reference counterpart
 * (the reference scans one s at a time); every s gets the line
 * gf2t_scan_run would give it, alone or in any mix.
 * One server per r.  Its slab holds a fixed number of slots, sized from
 * mem_bytes (0 = half of the free device memory): admission is by device
 * footprint.  Slot s values of any job advance together, m degrees per
 * round (the blocked scheme), each with its own degree frontier and
 * interval ends; queued s values enter free slots at round boundaries.
 * One caller at a time per server (calls take its lock).
 * ------------------------------------------------------------------ */
typedef struct gf2t_server gf2t_server;
int gf2t_server_create(uint64_t r, const gf2t_scan_params *params,
                       size_t mem_bytes, gf2t_server **out);
int gf2t_server_capacity(const gf2t_server *sv, size_t *slots,
                         size_t *bytes_per_slot);
/* queue the s values of job `job`: sieve survivors at depth k0, scanned
 * to min(maxd, Swan bound) (maxd 0 = the Swan bound) */
int gf2t_server_submit(gf2t_server *sv, uint64_t job, const uint64_t *svals,
                       size_t n, uint64_t k0, uint64_t maxd);
/* up to `rounds` rounds (admission, m degrees, interval checks); reports
 * the slots in use and the s values still queued */
int gf2t_server_step(gf2t_server *sv, int rounds, size_t *live,
                     size_t *queued);
/* finished lines since the last poll: up to max_rows rows, "\n"-separated
 * into buf (cap bytes), jobs[i] the job of row i, *nrows the count */
int gf2t_server_poll(gf2t_server *sv, size_t max_rows, uint64_t *jobs,
                     char *buf, size_t cap, size_t *nrows);
void gf2t_server_destroy(gf2t_server *sv);

/* Host gcd of two GF(2)[x] word arrays: degree (-1 if both zero) and up
 * to gcap words of the monic gcd.  Half-gcd recursion with Karatsuba
 * products, and device FFT products above the "gcd_dev_bits" threshold
 * (poly_gcd_naive / poly_gcd_ntl / poly_gcd_hybrid, src/gf2_gcd.h:48-75,
 * src/gf2_gcd_ntl.cpp:32-42, :354-452).  Returns GF2T_EHIP (negative)
 * if a device product fails. */
int64_t gf2t_gcd_host(const uint64_t *a, size_t aw, const uint64_t *b,
                      size_t bw, uint64_t *g, size_t gcap);
/* Numerically least irreducible factor of t, whose factors all have
 * degree d (edf_least_ntl / edf_least_enum, src/gf2_gcd_ntl.cpp:44-61,
 * src/gf2_gcd.h:137-145). */
int gf2t_edf_least(const uint64_t *t, size_t tw, uint64_t d, uint64_t *out,
                   size_t cap);
/* Degree bound of the scan: r/3 under Swan's criterion, else r/2
 * (src/tsfactor.cu:1101-1107); *swan (may be NULL) gets the criterion. */
uint64_t gf2t_swan_bound(uint64_t r, uint64_t s, int *swan);

#ifdef __cplusplus
}
#endif
#endif /* GF2T_H */
