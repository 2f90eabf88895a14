// gf2t_engine.hpp -- C++ engine structs with the reference's names and
// member semantics, implemented over the C ABI (gf2t.h), so code written
// against src/gf2_cantor_engine.h, src/gf2_cantor_cuda.h and
// src/gf2_trinomial_cuda.h recompiles against the MI355X library.
// Like the reference, a failed call prints the error and exits(1)
// (CUCHK, src/gf2_cantor_cuda.h:65-69); use gf2t.h directly for status
// returns.
#ifndef GF2T_ENGINE_HPP
#define GF2T_ENGINE_HPP
#include "gf2t.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace gf2t {

typedef uint64_t u64;
typedef uint32_t u32;

inline void check(int rc, const char *what) {
    if (rc != GF2T_OK) {
        fprintf(stderr, "gf2t error %d in %s: %s\n", rc, what, gf2t_last_error());
        exit(1);
    }
}

inline size_t bits_to_words(u64 bits) { return (size_t)((bits + 63) >> 6); }
inline size_t mul_out_words(u64 abits, u64 bbits) { return gf2t_mul_out_words(abits, bbits); }

// A device allocation of n words (RAII).
struct DevWords {
    u64 *p = nullptr;
    size_t n = 0;
    DevWords() {}
    explicit DevWords(size_t words) { alloc(words); }
    DevWords(const DevWords &) = delete;
    DevWords &operator=(const DevWords &) = delete;
    ~DevWords() { release(); }
    void alloc(size_t words) {
        release();
        void *q = nullptr;
        check(gf2t_malloc(&q, (words ? words : 1) * 8), "gf2t_malloc");
        p = (u64 *)q;
        n = words;
    }
    void release() {
        if (p) gf2t_free(p);
        p = nullptr;
        n = 0;
    }
    void upload(const u64 *h, size_t words) {
        check(gf2t_memcpy_h2d(p, h, words * 8, nullptr), "h2d");
    }
    void download(u64 *h, size_t words) const {
        check(gf2t_memcpy_d2h(h, p, words * 8, nullptr), "d2h");
    }
};

// GF(2^64) = GF(2)[x]/(x^64 + x^4 + x^3 + x + 1), bit-serial (host checks)
inline u64 gf64_mul_host(u64 a, u64 b) {
    u64 r = 0;
    while (b) {
        if (b & 1) r ^= a;
        u64 hi = a >> 63;
        a <<= 1;
        if (hi) a ^= 0x1BULL;
        b >>= 1;
    }
    return r;
}

// src/gf2_cantor_engine.h:74-121
struct CantorBasis {
    u64 b[64];
    int len = 0;
    void build() { check(gf2t_cantor_basis(b, &len), "gf2t_cantor_basis"); }
    bool verify() const {
        if (len < 1 || b[0] != 1) return false;
        for (int i = 1; i < len; i++)
            if ((gf64_mul_host(b[i], b[i]) ^ b[i]) != b[i - 1]) return false;
        u64 pv[64] = {0};
        for (int i = 0; i < len; i++) {
            u64 v = b[i];
            for (int bit = 63; bit >= 0 && v; bit--) {
                if (!((v >> bit) & 1)) continue;
                if (pv[bit]) v ^= pv[bit];
                else { pv[bit] = v; v = 0; }
            }
            if (v) return false;
            bool placed = false;
            for (int bit = 0; bit < 64; bit++) placed |= pv[bit] != 0;
            if (!placed) return false;
        }
        int rank = 0;
        for (int bit = 0; bit < 64; bit++) rank += pv[bit] != 0;
        return rank == len;
    }
};

// Device transform (CudaFFT, src/gf2_cantor_cuda.h:361-521).
struct CudaFFT {
    int m = 0;
    size_t n = 0;
    gf2t_fft_plan *plan = nullptr;
    void init(int m_, const CantorBasis &) {
        fini();
        m = m_;
        n = (size_t)1 << m;
        check(gf2t_fft_plan_create(m, &plan), "gf2t_fft_plan_create");
    }
    void fini() {
        if (plan) gf2t_fft_plan_destroy(plan);
        plan = nullptr;
    }
    void fwd(u64 *df) const { check(gf2t_fft_fwd(plan, df, nullptr), "gf2t_fft_fwd"); }
    void inv(u64 *df) const { check(gf2t_fft_inv(plan, df, nullptr), "gf2t_fft_inv"); }
};

// Host-buffer transform with the CantorFFT interface
// (src/gf2_cantor_engine.h:125-188); it runs on the device.
struct CantorFFT {
    int m = 0;
    size_t n = 0;
    CudaFFT dev;
    void init(int m_, const CantorBasis &B) {
        m = m_;
        n = (size_t)1 << m;
        dev.init(m, B);
    }
    void fwd(u64 *f) const { run(f, true); }
    void inv(u64 *f) const { run(f, false); }
    void run(u64 *f, bool forward) const {
        DevWords d(n);
        d.upload(f, n);
        if (forward) dev.fwd(d.p);
        else dev.inv(d.p);
        d.download(f, n);
    }
};

// out (mul_out_words words) = a * b over GF(2)[x] (gf2x_fft_mul,
// src/gf2_cantor_engine.h:232-242); host buffers in and out.
inline void gf2x_fft_mul(const u64 *a, u64 abits, const u64 *b, u64 bbits, u64 *out,
                         const CantorBasis &) {
    size_t ow = mul_out_words(abits, bbits);
    DevWords da(bits_to_words(abits)), db(bits_to_words(bbits)), dout(ow);
    da.upload(a, bits_to_words(abits));
    db.upload(b, bits_to_words(bbits));
    void *work = nullptr;
    size_t wb = gf2t_mul_workspace_bytes(abits, bbits);
    if (wb) check(gf2t_malloc(&work, wb), "gf2t_malloc");
    check(gf2t_mul(da.p, abits, db.p, bbits, dout.p, work, nullptr), "gf2t_mul");
    dout.download(out, ow);
    if (work) gf2t_free(work);
}

// Arithmetic in GF(2)[x]/(x^r + x^s + 1) (TrinGPU,
// src/gf2_trinomial_cuda.h:87-178).  dT0 holds the last full product
// (2r-1 bits) after mul_dev, as in the reference.
struct TrinGPU {
    u64 r = 0, s = 0;
    size_t nw = 0, nw2 = 0;
    int m = 0;
    CantorBasis Bb;
    DevWords A, ACC, T0, T1, W;
    u64 *dA = nullptr, *dACC = nullptr, *dT0 = nullptr, *dT1 = nullptr;

    void init(u64 r_) {
        r = r_;
        s = 0;
        nw = bits_to_words(r);
        nw2 = 2 * nw;
        Bb.build();
        size_t ca = (size_t)((r + 31) >> 5);
        m = 1;
        while (((size_t)1 << m) < 2 * ca - 1) m++;
        A.alloc(nw);
        ACC.alloc(nw);
        T0.alloc(nw2 + 2);
        T1.alloc(nw);
        W.alloc((gf2t_mul_workspace_bytes(r, r) + 7) / 8);
        dA = A.p; dACC = ACC.p; dT0 = T0.p; dT1 = T1.p;
    }
    void set_s(u64 s_) {
        if (!(s_ >= 1 && 2 * s_ < r)) {
            fprintf(stderr, "need 1 <= s <= (r-1)/2\n");
            exit(1);
        }
        s = s_;
    }
    void set_x() { check(gf2t_set_x(dA, nw, nullptr), "gf2t_set_x"); }
    void reset_acc() { check(gf2t_set_one(dACC, nw, nullptr), "gf2t_set_one"); }
    void reduce(u64 *dst_rbits, u64 L) {
        check(gf2t_fold(dT0, L, dst_rbits, r, s, nullptr), "gf2t_fold");
    }
    void square_A() {
        check(gf2t_square_fold(dA, dT1, r, s, nullptr), "gf2t_square_fold");
        std::swap(dA, dT1);
    }
    void mul_dev(const u64 *dx, const u64 *dy) {
        check(gf2t_mul(dx, r, dy, r, dT0, W.p, nullptr), "gf2t_mul");
    }
    void accumulate() {
        check(gf2t_xor_x(dA, nullptr), "gf2t_xor_x");
        check(gf2t_mulmod(dACC, r, dA, r, r, s, dACC, W.p, nullptr), "gf2t_mulmod");
        check(gf2t_xor_x(dA, nullptr), "gf2t_xor_x");
    }
    void fini() {
        A.release(); ACC.release(); T0.release(); T1.release(); W.release();
        dA = dACC = dT0 = dT1 = nullptr;
    }
};

// Device context of the blocked distinct-degree scan (TsGPU,
// src/tsfactor.cu:207-391): the h-table ping-pong dH[2] of m slots (slot
// j at dH[hcur] + j*hpitch; hpitch is nw rounded up to 32 words so that
// the two-squaring kernels can run on the table), the interval
// accumulator dA, the block polynomial dC and a 2r-bit product buffer
// dS0.  h_to_host / h_from_host exchange the table in the reference's
// packed layout (slot j at word j*nw), the h section of its TSFC
// checkpoint.  Every member has the reference's meaning:
//   init_h_at(k)      slot j = e_(j+1)(x^(2^k), ..., x^(2^(k+m-1))) mod T
//   square_all_step() every slot squared once (the table of block k+1)
//   horner_C()        dC = x^m + sum_j h_j x^(m-1-j) mod T
//                        = prod_t (x + x^(2^(k+t))) mod T
//   mul_full(x, y)    dS0 = x*y (2r-1 bits);  reduce_into(dst): dS0 mod T
//   mul_A_by_C()      dA = dA * dC mod T;     reset_A(): dA = 1
// legacy_sq is accepted: the reference's three-kernel squaring gives the
// same words as the single pass used here.
struct TsGPU {
    u64 r = 0, s = 0;
    size_t nw = 0, nw2 = 0, nwc = 0, hpitch = 0;
    int m = 0;
    int fftm = 0;
    CantorBasis Bb;
    DevWords H0, H1, Hold, A, Cw, Tmp, S0, S1, W, Sdev;
    u64 *dH[2] = {nullptr, nullptr};
    int hcur = 0;
    u64 *dHold = nullptr, *dA = nullptr, *dC = nullptr, *dTmp = nullptr, *dS0 = nullptr,
        *dS1 = nullptr;
    bool legacy_sq = false;
    double t_sq = 0, t_horner = 0, t_mul = 0;

    void init(u64 r_) {
        r = r_;
        nw = bits_to_words(r);
        nw2 = 2 * nw;
        hpitch = (nw + 31) & ~(size_t)31;
        Bb.build();
        if (!Bb.verify()) {
            fprintf(stderr, "Cantor basis failure\n");
            exit(1);
        }
        size_t ca = (size_t)((r + 31) >> 5);
        fftm = 1;
        while (((size_t)1 << fftm) < 2 * ca - 1) fftm++;
        A.alloc(nw);
        Cw.alloc(nw + 1);
        Tmp.alloc(nw);
        S0.alloc(mul_out_words(r, r));
        S1.alloc(nw2);
        W.alloc((gf2t_mul_workspace_bytes(r, r) + 7) / 8);
        Sdev.alloc(1);
        dA = A.p; dC = Cw.p; dTmp = Tmp.p; dS0 = S0.p; dS1 = S1.p;
    }
    void set_m(int m_) {
        if (m_ < 2 || m_ > 64) {
            fprintf(stderr, "need 2 <= m <= 64\n");
            exit(1);
        }
        if (m == m_) return;
        m = m_;
        nwc = bits_to_words(r + (u64)m - 1);
        H0.alloc((size_t)m * hpitch);
        H1.alloc((size_t)m * hpitch);
        Hold.release();
        dHold = nullptr;
        dH[0] = H0.p;
        dH[1] = H1.p;
        hcur = 0;
    }
    void need_hold() {
        if (!dHold) {
            Hold.alloc((size_t)m * hpitch);
            dHold = Hold.p;
        }
    }
    void set_s(u64 s_) {
        if (!(s_ >= 1 && 2 * s_ < r)) {
            fprintf(stderr, "need 1 <= s <= (r-1)/2 (got s=%llu)\n", (unsigned long long)s_);
            exit(1);
        }
        s = s_;
        Sdev.upload(&s, 1);
    }
    void fini() {
        H0.release(); H1.release(); Hold.release(); A.release(); Cw.release(); Tmp.release();
        S0.release(); S1.release(); W.release(); Sdev.release();
        dH[0] = dH[1] = dHold = dA = dC = dTmp = dS0 = dS1 = nullptr;
        m = 0;
    }
    u64 *hslot(int j) { return dH[hcur] + (size_t)j * hpitch; }
    u64 *hslot_o(int j) { return dH[1 - hcur] + (size_t)j * hpitch; }

    void square_slot(const u64 *src, u64 *dst) {
        check(gf2t_square_fold(src, dst, r, s, nullptr), "gf2t_square_fold");
    }
    void square_all_step() {
        check(gf2t_square_fold_slab_g(dH[hcur], dH[1 - hcur], hpitch, r, Sdev.p, (size_t)m, m,
                                      nullptr), "gf2t_square_fold_slab_g");
        hcur = 1 - hcur;
    }
    void horner_C() {
        check(gf2t_horner_slab(dH[hcur], hpitch, m, r, Sdev.p, dC, nw + 1, 1, nullptr),
              "gf2t_horner_slab");
    }
    void mul_full(const u64 *dx, const u64 *dy) {
        check(gf2t_mul(dx, r, dy, r, dS0, W.p, nullptr), "gf2t_mul");
    }
    void reduce_into(u64 *dst) {
        check(gf2t_fold(dS0, 2 * r - 1, dst, r, s, nullptr), "gf2t_fold");
    }
    void mul_A_by_C() {
        mul_full(dA, dC);
        reduce_into(dA);
    }
    void reset_A() { check(gf2t_set_one(dA, nw, nullptr), "gf2t_set_one"); }
    void init_h_at(u64 kstart) {
        check(gf2t_ddf_init_slab(dH[hcur], dH[1 - hcur], hpitch, m, r, Sdev.p, kstart, 1, nullptr),
              "gf2t_ddf_init_slab");
    }
    void copy_h(u64 *dst_base, const u64 *src_base) {
        check(gf2t_memcpy_d2d(dst_base, src_base, (size_t)m * hpitch * 8, nullptr), "d2d");
    }
    void h_to_host(std::vector<u64> &out) {
        std::vector<u64> tab((size_t)m * hpitch);
        check(gf2t_memcpy_d2h(tab.data(), dH[hcur], tab.size() * 8, nullptr), "d2h");
        out.resize((size_t)m * nw);
        for (int j = 0; j < m; j++)
            memcpy(out.data() + (size_t)j * nw, tab.data() + (size_t)j * hpitch, nw * 8);
    }
    void h_from_host(const std::vector<u64> &in) {
        std::vector<u64> tab((size_t)m * hpitch, 0);
        for (int j = 0; j < m; j++)
            memcpy(tab.data() + (size_t)j * hpitch, in.data() + (size_t)j * nw, nw * 8);
        check(gf2t_memcpy_h2d(dH[hcur], tab.data(), tab.size() * 8, nullptr), "h2d");
    }
    void A_to_host(std::vector<u64> &out) {
        out.resize(nw);
        check(gf2t_memcpy_d2h(out.data(), dA, nw * 8, nullptr), "d2h");
    }
    void A_from_host(const std::vector<u64> &in) {
        check(gf2t_memcpy_h2d(dA, in.data(), nw * 8, nullptr), "h2d");
    }
};

// Device product service with the member functions of the reference's
// GpuMulHook (src/gf2_gcd.h:155-172): mul() is the full product, and
// mat2_apply() the 2x2 matrix-vector product in one set of six forward
// and two inverse transforms.  Both return false (decline) on a library
// error, so a caller falls back to its own products.  Code that includes
// src/gf2_gcd.h binds it as
//   struct MI355XMulHook : GpuMulHook {
//       gf2t::MulService svc;
//       bool mul(const u64 *a, size_t aw, u64 abits, const u64 *b, size_t bw,
//                u64 bbits, std::vector<u64> &out) override {
//           return svc.mul(a, aw, abits, b, bw, bbits, out);
//       }
//       bool mat2_apply(const Op m[4], const Op &a, const Op &b,
//                       std::vector<u64> &oa, std::vector<u64> &ob) override {
//           gf2t::MulService::Op mm[4] = {{m[0].w, m[0].nw, m[0].bits}, ...};
//           return svc.mat2_apply(mm, {a.w, a.nw, a.bits}, {b.w, b.nw, b.bits}, oa, ob);
//       }
//   };
struct MulService {
    struct Op {
        const u64 *w;
        size_t nw;
        u64 bits;
    };
    bool mul(const u64 *a, size_t aw, u64 abits, const u64 *b, size_t bw, u64 bbits,
             std::vector<u64> &out) const {
        if (abits > 64 * (u64)aw || bbits > 64 * (u64)bw) return false;
        out.assign(mul_out_words(abits, bbits), 0);
        return gf2t_mul_host(a, abits, b, bbits, out.data()) == GF2T_OK;
    }
    bool mat2_apply(const Op m[4], const Op &a, const Op &b, std::vector<u64> &oa,
                    std::vector<u64> &ob) const {
        const uint64_t *hm[4];
        uint64_t mb[4], maxm = 0;
        for (int i = 0; i < 4; i++) {
            if (m[i].bits > 64 * (u64)m[i].nw) return false;
            hm[i] = m[i].w;
            mb[i] = m[i].bits;
            if (mb[i] > maxm) maxm = mb[i];
        }
        if (a.bits > 64 * (u64)a.nw || b.bits > 64 * (u64)b.nw) return false;
        u64 maxv = a.bits > b.bits ? a.bits : b.bits;
        size_t ow = mul_out_words(maxm, maxv);
        oa.assign(ow, 0);
        ob.assign(ow, 0);
        return gf2t_mat2_host(hm, mb, a.w, a.bits, b.w, b.bits, oa.data(), ob.data(), ow) ==
               GF2T_OK;
    }
};

// Coarse sieve of all s in [1, smax] over degrees min_depth..depth into a
// host best[] (run_gpu, src/coarse_sieve.cu:1119-1225); best seeds the
// minimum when it already holds smax+1 entries.
inline void run_gpu(u64 r, int depth, u64 smax, std::vector<u64> &best, int min_depth = 2,
                    std::vector<gf2t_sieve_stats> *stats = nullptr) {
    if (best.size() != smax + 1) best.assign(smax + 1, ~0ULL);
    DevWords d(smax + 1);
    d.upload(best.data(), smax + 1);
    std::vector<gf2t_sieve_stats> st(37);
    check(gf2t_sieve(r, min_depth, depth, smax, d.p, st.data()), "gf2t_sieve");
    d.download(best.data(), smax + 1);
    if (stats) *stats = st;
}

}  // namespace gf2t
#endif
