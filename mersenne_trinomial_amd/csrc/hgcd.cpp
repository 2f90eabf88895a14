// hgcd.cpp -- subquadratic gcd in GF(2)[x] (half-gcd recursion), the role
// of the reference's poly_gcd_ntl / poly_gcd_hybrid (src/gf2_gcd_ntl.cpp:
// 32-42, :354-452): products below a size threshold run as Karatsuba over
// PCLMULQDQ on the calling host thread, products above it on the device
// through the Cantor-FFT product service (gf2t_mul_host, gf2t_mat2_host).
//
// Remainder sequence of (a, b), deg a > deg b:  r0 = a, r1 = b,
// r(i+1) = r(i-1) + q(i) r(i).  One step is the matrix [[0,1],[1,q]]
// acting on the column (r(i-1), r(i)); a product of steps M maps (a, b) to
// a pair of consecutive remainders.  hgcd(a, b) returns the M that reaches
// the pair (c, d) with deg c >= ceil(deg a / 2) > deg d.  Its first half
// recurses on the top coefficients of (a, b) only (the quotients while
// the degree stays above that half depend on nothing else), applies the
// result to the full pair, takes one explicit division step and recurses
// once more on the top of the new pair.
#include "host_poly.hpp"
#include "gf2t.h"
#include <immintrin.h>
#include <wmmintrin.h>
#include <smmintrin.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <future>
#include <mutex>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <sched.h>

namespace gf2t_host {

namespace {

std::atomic<u64> g_dev_bits{0};   // 0: host products only

// Per-thread phase times of the gcd this thread runs, printed to stderr
// at its end when GF2T_GCD_STATS is set (for gcds of at least 2^20 bits):
// calls into the device product service and the time the thread waited
// in them, and the word-level base case.  The rest of a gcd's time is
// host products, divisions and bookkeeping.
struct GcdStats {
    long long dev_calls = 0, dev_ns = 0, base_calls = 0, base_ns = 0;
};
thread_local GcdStats t_stats;
bool gcd_stats_on() {
    static const bool on = [] {
        const char *e = getenv("GF2T_GCD_STATS");
        return e && *e && *e != '0';
    }();
    return on;
}
inline long long now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}
// times one call for t_stats when statistics are on
struct PhaseTimer {
    long long *calls, *ns, t0;
    PhaseTimer(long long &c, long long &n) : calls(&c), ns(&n), t0(gcd_stats_on() ? now_ns() : 0) {}
    ~PhaseTimer() {
        if (!t0) return;
        ++*calls;
        *ns += now_ns() - t0;
    }
};

// Helper threads for independent products inside one gcd: a pool of
// persistent workers (started on demand, at most host threads - 1).  A
// task goes to the pool only while the threads computing (callers inside
// gcd_fast plus helpers with a task) leave a hardware thread free; else it
// runs on the calling thread.  So one gcd alone spreads over the host,
// while many concurrent gcds (one per thread) run without helpers and
// without oversubscribing the cores.  Every unfinished task counts in
// `busy` and there are at least `busy` workers, so a task waiting on a
// nested task always has a worker for it.
struct HelperPool {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::packaged_task<void()>> q;
    int workers = 0;
    std::atomic<int> busy{0};
};
HelperPool &helper_pool() {
    static HelperPool *p = new HelperPool;   // never destroyed: workers outlive main()
    return *p;
}
std::atomic<int> g_gcd_callers{0};   // threads inside gcd_fast

void helper_loop(HelperPool *P) {
    for (;;) {
        std::packaged_task<void()> task;
        {
            std::unique_lock<std::mutex> lk(P->mu);
            P->cv.wait(lk, [&] { return !P->q.empty(); });
            task = std::move(P->q.front());
            P->q.pop_front();
        }
        task();
        P->busy.fetch_sub(1);
    }
}

// run f on a helper when the budget allows, else now
template <class F>
std::future<void> spawn(F &&f) {
    HelperPool &P = helper_pool();
    const int cap = host_threads();
    int b = P.busy.load();
    while (g_gcd_callers.load(std::memory_order_relaxed) + b < cap) {
        if (!P.busy.compare_exchange_weak(b, b + 1)) continue;
        std::packaged_task<void()> task(std::forward<F>(f));
        std::future<void> fut = task.get_future();
        {
            std::lock_guard<std::mutex> lk(P.mu);
            P.q.push_back(std::move(task));
            if (P.workers < P.busy.load()) {
                std::thread(helper_loop, &P).detach();
                P.workers++;
            }
        }
        P.cv.notify_one();
        return fut;
    }
    f();
    std::promise<void> p;
    p.set_value();
    return p.get_future();
}
constexpr size_t PAR_WORDS = 512;   // products with both operands this large may go parallel

// words: Karatsuba's schoolbook base (gf2t_tune "gcd_school_words")
std::atomic<size_t> g_kbase{48};
// bits: the word-window half-gcd below this (gf2t_tune "gcd_base_bits")
std::atomic<int64_t> g_hgcd_base{3072};
constexpr int64_t GCD_BASE = 4096;     // bits: plain Euclid below this

inline void clmul(u64 a, u64 b, u64 &lo, u64 &hi) {
    __m128i p = _mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a),
                                     _mm_cvtsi64_si128((long long)b), 0);
    lo = (u64)_mm_cvtsi128_si64(p);
    hi = (u64)_mm_extract_epi64(p, 1);
}

// out[0, na+nb) ^= a[0,na) * b[0,nb), column by column: column k sums the
// 128-bit products a[i] b[k-i] in registers (two per pair of 16-byte
// loads, into two accumulators) and passes its high half to column k+1,
// so each output word is written once
void school128(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out) {
    if (!na || !nb) return;
    const size_t nk = na + nb - 1;
    u64 carry = 0;
    for (size_t k = 0; k < nk; k++) {
        const size_t i0 = k >= nb ? k - nb + 1 : 0;
        const size_t i1 = k < na ? k : na - 1;        // last i, inclusive
        __m128i s0 = _mm_setzero_si128(), s1 = _mm_setzero_si128();
        size_t i = i0;
        for (; i < i1; i += 2) {
            // (a[i], a[i+1]) against (b[k-i-1], b[k-i])
            const __m128i A = _mm_loadu_si128((const __m128i *)(a + i));
            const __m128i B = _mm_loadu_si128((const __m128i *)(b + (k - i - 1)));
            s0 = _mm_xor_si128(s0, _mm_clmulepi64_si128(A, B, 0x10));
            s1 = _mm_xor_si128(s1, _mm_clmulepi64_si128(A, B, 0x01));
        }
        if (i == i1)
            s0 = _mm_xor_si128(s0, _mm_clmulepi64_si128(_mm_loadl_epi64((const __m128i *)(a + i)),
                                                       _mm_loadl_epi64((const __m128i *)(b + (k - i))),
                                                       0x00));
        s0 = _mm_xor_si128(s0, s1);
        out[k] ^= (u64)_mm_cvtsi128_si64(s0) ^ carry;
        carry = (u64)_mm_extract_epi64(s0, 1);
    }
    out[nk] ^= carry;
}

// out[0, na+nb) ^= a * b with 512-bit VPCLMULQDQ, by output blocks of
// eight words: for block k and each word b[j] that reaches it, the eight
// words a[k-j .. k-j+8) (from a copy of a with eight zero words on each
// side) are multiplied by b[j] in two instructions, the even words' 128-bit
// products landing on output words (k+2l, k+2l+1) and the odd words' one
// word higher; the two accumulators are combined once per block, the odd
// one shifted up by a word with its top word carried into the next block.
__attribute__((target("avx512f,vpclmulqdq")))
void rows512(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out) {
    thread_local std::vector<u64> pad;   // grows to the largest na seen
    if (pad.size() < na + 24) pad.resize(na + 24);
    u64 *const ap = pad.data();
    std::memset(ap, 0, 8 * 8);
    std::memcpy(ap + 8, a, na * 8);
    std::memset(ap + 8 + na, 0, 16 * 8);
    const size_t no = na + nb;          // output words
    __m512i prevO = _mm512_setzero_si512();
    for (size_t k = 0; k < no; k += 8) {
        const size_t jlo = k + 1 > na ? k + 1 - na : 0;
        const size_t jhi = std::min(nb, k + 8);          // exclusive
        __m512i E0 = _mm512_setzero_si512(), O0 = _mm512_setzero_si512();
        __m512i E1 = _mm512_setzero_si512(), O1 = _mm512_setzero_si512();
        size_t j = jlo;
        for (; j + 2 <= jhi; j += 2) {
            const __m512i A0 = _mm512_loadu_si512(ap + 8 + k - j);
            const __m512i B0 = _mm512_set1_epi64((long long)b[j]);
            const __m512i A1 = _mm512_loadu_si512(ap + 7 + k - j);
            const __m512i B1 = _mm512_set1_epi64((long long)b[j + 1]);
            E0 = _mm512_xor_si512(E0, _mm512_clmulepi64_epi128(A0, B0, 0x00));
            O0 = _mm512_xor_si512(O0, _mm512_clmulepi64_epi128(A0, B0, 0x01));
            E1 = _mm512_xor_si512(E1, _mm512_clmulepi64_epi128(A1, B1, 0x00));
            O1 = _mm512_xor_si512(O1, _mm512_clmulepi64_epi128(A1, B1, 0x01));
        }
        if (j < jhi) {
            const __m512i A0 = _mm512_loadu_si512(ap + 8 + k - j);
            const __m512i B0 = _mm512_set1_epi64((long long)b[j]);
            E0 = _mm512_xor_si512(E0, _mm512_clmulepi64_epi128(A0, B0, 0x00));
            O0 = _mm512_xor_si512(O0, _mm512_clmulepi64_epi128(A0, B0, 0x01));
        }
        const __m512i E = _mm512_xor_si512(E0, E1), O = _mm512_xor_si512(O0, O1);
        const __m512i r = _mm512_xor_si512(E, _mm512_alignr_epi64(O, prevO, 7));
        prevO = O;
        if (k + 8 <= no) {
            _mm512_storeu_si512(out + k, _mm512_xor_si512(_mm512_loadu_si512(out + k), r));
        } else {
            const __mmask8 m = (__mmask8)((1u << (no - k)) - 1);
            _mm512_mask_storeu_epi64(out + k, m, _mm512_xor_si512(_mm512_maskz_loadu_epi64(m, out + k), r));
        }
    }
}

// The 512-bit form where the CPU has it (AVX-512F and VPCLMULQDQ), unless
// GF2T_NO_VPCLMUL is set
bool use512() {
    static const bool yes = [] {
        __builtin_cpu_init();
        const char *e = getenv("GF2T_NO_VPCLMUL");
        return !(e && *e && *e != '0') && __builtin_cpu_supports("avx512f") &&
               __builtin_cpu_supports("vpclmulqdq");
    }();
    return yes;
}

void school(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out) {
    if (!na || !nb) return;
    if (na < nb) {
        std::swap(a, b);
        std::swap(na, nb);
    }
    if (use512())
        rows512(a, na, b, nb, out);
    else
        school128(a, na, b, nb, out);
}

// Karatsuba's recombination: with z0 = a0 b0 in out[0, 2h), z2 = a1 b1 in
// out[2h, 2h+2l) and z1 = (a0+a1)(b0+b1), out[h, 3h) ^= z1 + z0 + z2.
// Step i reads the four words i, h+i, 2h+i, 3h+i of out and z1[i],
// z1[h+i], and writes out[h+i] and out[2h+i], which no later step reads.
__attribute__((target("avx512f")))
void kara_mix512(u64 *out, const u64 *z1, size_t h, size_t l) {
    const size_t nd = 2 * l - h;                  // steps with a z2[h+i] word
    size_t i = 0;
    for (; i + 8 <= h; i += 8) {
        const __m512i A = _mm512_loadu_si512(out + i), B = _mm512_loadu_si512(out + h + i);
        const __m512i Cv = _mm512_loadu_si512(out + 2 * h + i);
        __m512i D = _mm512_setzero_si512();
        if (i + 8 <= nd)
            D = _mm512_loadu_si512(out + 3 * h + i);
        else if (i < nd)
            D = _mm512_maskz_loadu_epi64((__mmask8)((1u << (nd - i)) - 1), out + 3 * h + i);
        const __m512i lo = _mm512_xor_si512(_mm512_xor_si512(_mm512_loadu_si512(z1 + i), A), Cv);
        const __m512i hi = _mm512_xor_si512(_mm512_xor_si512(_mm512_loadu_si512(z1 + h + i), B), D);
        _mm512_storeu_si512(out + h + i, _mm512_xor_si512(B, lo));
        _mm512_storeu_si512(out + 2 * h + i, _mm512_xor_si512(Cv, hi));
    }
    for (; i < h; i++) {
        const u64 A = out[i], B = out[h + i], Cv = out[2 * h + i], D = i < nd ? out[3 * h + i] : 0;
        out[h + i] = B ^ z1[i] ^ A ^ Cv;
        out[2 * h + i] = Cv ^ z1[h + i] ^ B ^ D;
    }
}

void kara_mix(u64 *out, const u64 *z1, size_t h, size_t l) {
    if (use512()) return kara_mix512(out, z1, h, l);
    const size_t nd = 2 * l - h;
    for (size_t i = 0; i < h; i++) {
        const u64 A = out[i], B = out[h + i], Cv = out[2 * h + i], D = i < nd ? out[3 * h + i] : 0;
        out[h + i] = B ^ z1[i] ^ A ^ Cv;
        out[2 * h + i] = Cv ^ z1[h + i] ^ B ^ D;
    }
}

// out[0, 2n) = a[0,n) * b[0,n); ws holds at least 4n words per level
void kara(const u64 *a, const u64 *b, size_t n, u64 *out, u64 *ws) {
    if (n <= g_kbase.load(std::memory_order_relaxed)) {
        std::memset(out, 0, 2 * n * 8);
        school(a, n, b, n, out);
        return;
    }
    size_t h = (n + 1) / 2, l = n - h;
    u64 *sa = ws, *sb = ws + h, *z1 = ws + 2 * h, *next = ws + 4 * h;
    for (size_t i = 0; i < l; i++) {
        sa[i] = a[i] ^ a[h + i];
        sb[i] = b[i] ^ b[h + i];
    }
    if (l < h) {
        sa[h - 1] = a[h - 1];
        sb[h - 1] = b[h - 1];
    }
    if (n >= 2 * PAR_WORDS) {
        // the three half products are independent; each helper gets its
        // own workspace (left uninitialized)
        std::unique_ptr<u64[]> w0(new u64[8 * h + 64]), w2(new u64[8 * h + 64]);
        auto f0 = spawn([&] { kara(a, b, h, out, w0.get()); });
        auto f2 = spawn([&] { kara(a + h, b + h, l, out + 2 * h, w2.get()); });
        kara(sa, sb, h, z1, next);
        f0.get();
        f2.get();
    } else {
        kara(a, b, h, out, next);                 // z0 -> out[0, 2h)
        kara(a + h, b + h, l, out + 2 * h, next); // z2 -> out[2h, 2n)
        kara(sa, sb, h, z1, next);                // z1 (2h words)
    }
    kara_mix(out, z1, h, l);
}

size_t words(const Poly &a) {
    int64_t d = deg(a);
    return d < 0 ? 0 : (size_t)(d >> 6) + 1;
}

void check_rc(int rc, const char *what) {
    if (rc != GF2T_OK)
        throw std::runtime_error(std::string(what) + ": " + gf2t_last_error());
}

Poly host_mul(const u64 *a, size_t na, const u64 *b, size_t nb) {
    if (na < nb) {
        std::swap(a, b);
        std::swap(na, nb);
    }
    Poly out(na + nb + 1, 0);
    if (nb <= g_kbase.load(std::memory_order_relaxed)) {
        school(a, na, b, nb, out.data());
        trim(out);
        return out;
    }
    // a in blocks of nb words, each block times b by Karatsuba (scratch
    // left uninitialized: every word is written before it is read)
    std::unique_ptr<u64[]> scratch(new u64[8 * nb + 64 + 3 * nb]);
    u64 *const wsp = scratch.get(), *const blk = wsp + 8 * nb + 64, *const part = blk + nb;
    for (size_t off = 0; off < na; off += nb) {
        size_t len = std::min(nb, na - off);
        std::memcpy(blk, a + off, len * 8);
        if (len < nb) std::memset(blk + len, 0, (nb - len) * 8);
        kara(blk, b, nb, part, wsp);
        size_t lim = std::min(2 * nb, out.size() - off);
        for (size_t i = 0; i < lim; i++) out[off + i] ^= part[i];
    }
    trim(out);
    return out;
}

Poly add(const Poly &a, const Poly &b) {
    Poly out = a.size() >= b.size() ? a : b;
    const Poly &o = a.size() >= b.size() ? b : a;
    for (size_t i = 0; i < o.size(); i++) out[i] ^= o[i];
    trim(out);
    return out;
}

// a div x^k
Poly shr(const Poly &a, u64 k) {
    size_t ws = (size_t)(k >> 6), n = words(a);
    unsigned bs = (unsigned)(k & 63);
    if (ws >= n) return Poly();
    Poly out(n - ws);
    for (size_t i = 0; i < out.size(); i++) {
        u64 lo = a[ws + i] >> bs;
        u64 hi = (bs && ws + i + 1 < n) ? a[ws + i + 1] << (64 - bs) : 0;
        out[i] = lo | hi;
    }
    trim(out);
    return out;
}

// a mod x^k
Poly low(const Poly &a, u64 k) {
    size_t n = std::min(words(a), (size_t)((k + 63) >> 6));
    Poly out(a.begin(), a.begin() + n);
    if (n && (k & 63) && n == (size_t)((k + 63) >> 6)) out[n - 1] &= (1ULL << (k & 63)) - 1;
    trim(out);
    return out;
}

void divmod(const Poly &a, const Poly &b, Poly &q, Poly &r) {
    int64_t db = deg(b), dr = deg(a);
    if (db < 0) throw std::runtime_error("gcd: division by zero");
    r = a;
    r.resize(std::max(r.size(), words(b)) + 2, 0);
    q.assign(dr >= db ? (size_t)((dr - db) >> 6) + 1 : 1, 0);
    size_t bw = words(b);
    while (dr >= db) {
        u64 sh = (u64)(dr - db);
        q[sh >> 6] |= 1ULL << (sh & 63);
        xor_shift_words(r.data(), b.data(), bw, sh);
        dr = deg_w(r.data(), (size_t)(dr >> 6) + 1);
    }
    trim(q);
    trim(r);
}

// (x, y) -> (M00 x + M01 y, M10 x + M11 y)
void apply(const HMat &M, const Poly &x, const Poly &y, Poly &ox, Poly &oy) {
    int64_t dm = std::max(std::max(deg(M.m00), deg(M.m01)), std::max(deg(M.m10), deg(M.m11)));
    int64_t dv = std::max(deg(x), deg(y));
    u64 lim = g_dev_bits.load(std::memory_order_relaxed);
    if (lim && dm + 1 >= (int64_t)lim && dv + 1 >= (int64_t)lim) {
        const Poly *e[4] = {&M.m00, &M.m01, &M.m10, &M.m11};
        const uint64_t *hm[4];
        uint64_t mb[4];
        for (int i = 0; i < 4; i++) {
            hm[i] = e[i]->data();
            mb[i] = (uint64_t)(deg(*e[i]) + 1);
        }
        size_t ow = gf2t_mul_out_words((u64)dm + 1, (u64)dv + 1);
        ox.assign(ow, 0);
        oy.assign(ow, 0);
        PhaseTimer pt(t_stats.dev_calls, t_stats.dev_ns);
        check_rc(gf2t_mat2_host(hm, mb, x.data(), (u64)(deg(x) + 1), y.data(),
                                (u64)(deg(y) + 1), ox.data(), oy.data(), ow),
                 "gf2t_mat2_host");
        trim(ox);
        trim(oy);
        return;
    }
    if (std::min<int64_t>(dm, dv) >= 64 * (int64_t)PAR_WORDS) {
        Poly p0, p1, p2, p3;
        auto f1 = spawn([&] { p1 = mul_fast(M.m01, y); });
        auto f2 = spawn([&] { p2 = mul_fast(M.m10, x); });
        auto f3 = spawn([&] { p3 = mul_fast(M.m11, y); });
        p0 = mul_fast(M.m00, x);
        f1.get();
        f2.get();
        f3.get();
        ox = add(p0, p1);
        oy = add(p2, p3);
        return;
    }
    ox = add(mul_fast(M.m00, x), mul_fast(M.m01, y));
    oy = add(mul_fast(M.m10, x), mul_fast(M.m11, y));
}

// (ox_i, oy_i) = M (x_i, y_i) for i < n (n <= 3) with one set of matrix
// transforms on the device when the largest pair qualifies for it, else
// pair by pair
void applyN(const HMat &M, int n, const Poly *const *x, const Poly *const *y, Poly *const *ox,
            Poly *const *oy) {
    int64_t dm = std::max(std::max(deg(M.m00), deg(M.m01)), std::max(deg(M.m10), deg(M.m11)));
    int64_t dv = -1;
    for (int i = 0; i < n; i++) dv = std::max(dv, std::max(deg(*x[i]), deg(*y[i])));
    u64 lim = g_dev_bits.load(std::memory_order_relaxed);
    if (n > 1 && lim && dm + 1 >= (int64_t)lim && dv + 1 >= (int64_t)lim) {
        const Poly *e[4] = {&M.m00, &M.m01, &M.m10, &M.m11};
        const uint64_t *hm[4];
        uint64_t mb[4];
        for (int i = 0; i < 4; i++) {
            hm[i] = e[i]->data();
            mb[i] = (uint64_t)(deg(*e[i]) + 1);
        }
        const uint64_t *hv[6];
        uint64_t vb[6];
        uint64_t *ho[6];
        size_t ow = gf2t_mul_out_words((u64)dm + 1, (u64)dv + 1);
        for (int i = 0; i < n; i++) {
            hv[2 * i] = x[i]->data();
            vb[2 * i] = (uint64_t)(deg(*x[i]) + 1);
            hv[2 * i + 1] = y[i]->data();
            vb[2 * i + 1] = (uint64_t)(deg(*y[i]) + 1);
            ox[i]->assign(ow, 0);
            oy[i]->assign(ow, 0);
            ho[2 * i] = ox[i]->data();
            ho[2 * i + 1] = oy[i]->data();
        }
        {
            PhaseTimer pt(t_stats.dev_calls, t_stats.dev_ns);
            check_rc(gf2t_mat2n_host(hm, mb, n, hv, vb, ho, ow), "gf2t_mat2n_host");
        }
        for (int i = 0; i < n; i++) {
            trim(*ox[i]);
            trim(*oy[i]);
        }
        return;
    }
    for (int i = 0; i < n; i++) apply(M, *x[i], *y[i], *ox[i], *oy[i]);
}

// ---- word-level base case --------------------------------------------
inline int deg64(u64 v) { return v ? 63 - __builtin_clzll(v) : -1; }

inline u64 clmul_lo(u64 a, u64 b) {
    u64 lo, hi;
    clmul(a, b, lo, hi);
    return lo;
}

typedef unsigned __int128 u128;

inline int deg128(u128 v) {
    const u64 hi = (u64)(v >> 64);
    return hi ? 127 - __builtin_clzll(hi) : deg64((u64)v);
}

// bits [p, p + 128) of a (zero beyond its words)
inline u128 window128(const Poly &a, u64 p) {
    const size_t w = (size_t)(p >> 6), n = a.size();
    const unsigned off = (unsigned)(p & 63);
    const u64 w0 = w < n ? a[w] : 0, w1 = w + 1 < n ? a[w + 1] : 0, w2 = w + 2 < n ? a[w + 2] : 0;
    if (!off) return ((u128)w1 << 64) | w0;
    return ((u128)((w1 >> off) | (w2 << (64 - off))) << 64) | ((w0 >> off) | (w1 << (64 - off)));
}

struct W2 {
    u64 m00, m01, m10, m11;
};

// Remainder steps of a pair of at most 128-bit polynomials (deg c > deg d)
// while deg d >= stop; returns the step matrix.  Each entry has degree
// deg(c0) - deg(current c) <= 127 - stop, and each quotient at most
// deg(c) - stop, so both fit in one word when stop >= 64 (or deg c0 <=
// 63 + stop).
// the steps of a pair of at most 64-bit words (deg c > deg d) while deg d
// >= stop, applied to the step matrix rows as in hgcd_word below
inline W2 hgcd_word64(u64 c, u64 d, int stop, int &steps) {
    u64 m00 = 1, m01 = 0, m10 = 0, m11 = 1;
    steps = 0;
    int dc = deg64(c), dd = deg64(d);
    while (dd >= stop && dd >= 0) {
        u64 r = c, n10 = m00, n11 = m01;
        int dr = dc;
        while (dr >= dd) {
            const int k = dr - dd;
            r ^= d << k;
            n10 ^= m10 << k;
            n11 ^= m11 << k;
            dr = deg64(r);
        }
        m00 = m10;
        m01 = m11;
        m10 = n10;
        m11 = n11;
        c = d;
        dc = dd;
        d = r;
        dd = dr;
        steps++;
    }
    return W2{m00, m01, m10, m11};
}

// low 128 bits of the carry-less product of a word and a 128-bit value
inline u128 clmul_w128(u64 w, u128 v) {
    u64 l0, h0, l1, h1;
    clmul(w, (u64)v, l0, h0);
    clmul(w, (u64)(v >> 64), l1, h1);
    (void)h1;
    return ((u128)(h0 ^ l1) << 64) | l0;
}
// the product of two step matrices A B (entries stay single words: the
// product is itself a step matrix of at most 63 degrees)
inline W2 w2_mul(const W2 &A, const W2 &B) {
    return W2{clmul_lo(A.m00, B.m00) ^ clmul_lo(A.m01, B.m10),
              clmul_lo(A.m00, B.m01) ^ clmul_lo(A.m01, B.m11),
              clmul_lo(A.m10, B.m00) ^ clmul_lo(A.m11, B.m10),
              clmul_lo(A.m10, B.m01) ^ clmul_lo(A.m11, B.m11)};
}

// Remainder steps of a pair of at most 128-bit polynomials (deg c > deg d)
// while deg d >= stop; returns the step matrix.  Each entry has degree
// deg(c0) - deg(current c) <= 127 - stop, and each quotient at most
// deg(c) - stop, so both fit in one word when stop >= 64 (or deg c0 <=
// 63 + stop).  The steps run on the top 64 bits of the pair where they
// can (a sub-round takes the steps of those bits while their remainders
// keep half the 64-bit window, the rule of the recursion), and the
// sub-round's matrix is applied to the 128-bit pair by carry-less
// multiplies; the steps left run on 128 bits, the quotient applied to the
// matrix rows one set bit at a time.
W2 hgcd_word(u128 c, u128 d, int stop, int &steps) {
    W2 W{1, 0, 0, 1};
    steps = 0;
    int dc = deg128(c), dd = deg128(d);
    for (;;) {
        if (dd < stop || dd < 0) return W;
        if (dc < 64 + 32) break;   // too few steps left for a 64-bit sub-round
        const int sh = dc - 63;
        const int st64 = std::max(32, stop - sh);
        int n = 0;
        const W2 S = hgcd_word64((u64)(c >> sh), (u64)(d >> sh), st64, n);
        if (n == 0) break;
        const u128 nc = clmul_w128(S.m00, c) ^ clmul_w128(S.m01, d);
        const u128 nd = clmul_w128(S.m10, c) ^ clmul_w128(S.m11, d);
        c = nc;
        d = nd;
        W = w2_mul(S, W);
        steps += n;
        dc = deg128(c);
        dd = deg128(d);
    }
    u64 m00 = W.m00, m01 = W.m01, m10 = W.m10, m11 = W.m11;
    while (dd >= stop && dd >= 0) {
        u128 r = c;
        u64 n10 = m00, n11 = m01;
        int dr = dc;
        while (dr >= dd) {
            const int k = dr - dd;
            r ^= d << k;
            n10 ^= m10 << k;
            n11 ^= m11 << k;
            dr = deg128(r);
        }
        m00 = m10;
        m01 = m11;
        m10 = n10;
        m11 = n11;
        c = d;
        dc = dd;
        d = r;
        dd = dr;
        steps++;
    }
    return W2{m00, m01, m10, m11};
}

// lin2 with 512-bit VPCLMULQDQ, eight words per step: the even words'
// 128-bit products land on output words (i+2l, i+2l+1) and the odd
// words' one word higher, the odd accumulator shifted up by a word with
// its top word carried into the next step (as in rows512).  Words past
// x's or y's end load as zero; out gets ceil(n/8)*8 + 1 words.
__attribute__((target("avx512f,vpclmulqdq")))
void lin2_512(u64 w0, const u64 *x, size_t nx, u64 w1, const u64 *y, size_t ny, u64 *out,
              size_t n) {
    const __m512i W0 = _mm512_set1_epi64((long long)w0), W1 = _mm512_set1_epi64((long long)w1);
    auto mask = [](size_t len, size_t i) -> __mmask8 {
        return i >= len ? 0 : len - i >= 8 ? (__mmask8)0xFF : (__mmask8)((1u << (len - i)) - 1);
    };
    __m512i prevO = _mm512_setzero_si512();
    size_t i = 0;
    for (; i < n; i += 8) {
        const __m512i X = _mm512_maskz_loadu_epi64(mask(nx, i), x + i);
        const __m512i Y = _mm512_maskz_loadu_epi64(mask(ny, i), y + i);
        const __m512i E = _mm512_xor_si512(_mm512_clmulepi64_epi128(X, W0, 0x00),
                                           _mm512_clmulepi64_epi128(Y, W1, 0x00));
        const __m512i O = _mm512_xor_si512(_mm512_clmulepi64_epi128(X, W0, 0x01),
                                           _mm512_clmulepi64_epi128(Y, W1, 0x01));
        _mm512_storeu_si512(out + i, _mm512_xor_si512(E, _mm512_alignr_epi64(O, prevO, 7)));
        prevO = O;
    }
    out[i] = (u64)_mm256_extract_epi64(_mm512_extracti64x4_epi64(prevO, 1), 3);
}

// out = w0 x + w1 y (word polynomials times multiword ones), one pass:
// word i is lo(w0 x[i] + w1 y[i]) plus the previous word's high half
void lin2(u64 w0, const Poly &x, u64 w1, const Poly &y, Poly &out) {
    const size_t nx = x.size(), ny = y.size(), n = std::max(nx, ny);
    if (n >= 8 && use512()) {
        out.resize(((n + 7) & ~(size_t)7) + 1);
        lin2_512(w0, x.data(), nx, w1, y.data(), ny, out.data(), n);
        trim(out);
        return;
    }
    out.resize(n + 1);
    const __m128i W = _mm_set_epi64x((long long)w1, (long long)w0);
    u64 carry = 0;
    size_t i = 0;
    for (const size_t nb = std::min(nx, ny); i < nb; i++) {
        const __m128i X = _mm_set_epi64x((long long)y[i], (long long)x[i]);
        const __m128i p = _mm_xor_si128(_mm_clmulepi64_si128(W, X, 0x00),
                                        _mm_clmulepi64_si128(W, X, 0x11));
        out[i] = (u64)_mm_cvtsi128_si64(p) ^ carry;
        carry = (u64)_mm_extract_epi64(p, 1);
    }
    const Poly &z = nx > ny ? x : y;
    const __m128i Z = _mm_cvtsi64_si128((long long)(nx > ny ? w0 : w1));
    for (; i < n; i++) {
        const __m128i p = _mm_clmulepi64_si128(Z, _mm_cvtsi64_si128((long long)z[i]), 0x00);
        out[i] = (u64)_mm_cvtsi128_si64(p) ^ carry;
        carry = (u64)_mm_extract_epi64(p, 1);
    }
    out[n] = carry;
    trim(out);
}

// half-gcd for small inputs (deg a < the "gcd_base_bits" knob): rounds of half-gcds of
// the top 128 coefficients.  A window's quotients are those of the full
// pair while the window remainders keep at least half the window's degree
// (the same rule as the recursion), so each round applies a correct prefix
// of the remainder sequence, about 64 degrees long, whose matrix has
// single-word entries.  A round that cannot step (a quotient longer than
// the window) takes one explicit division.
void hgcd_base(const Poly &a, const Poly &b, int64_t m, HMat &M, Poly &c, Poly &d) {
    M = HMat{Poly{1}, Poly(), Poly(), Poly{1}};
    c = a;
    d = b;
    trim(c);
    trim(d);
    Poly t0, t1, t2, t3, q, r;
    while (deg(d) >= m) {
        const int64_t dc = deg(c);
        const int64_t sh = std::max<int64_t>(0, dc - 127);
        const u128 cw = window128(c, (u64)sh), dw = window128(d, (u64)sh);
        const int stop = sh == 0 ? (int)m
                                 : (int)std::max<int64_t>((dc - sh + 1) / 2, m - sh);
        int steps = 0;
        W2 W = stop <= 127 ? hgcd_word(cw, dw, std::max(stop, 0), steps) : W2{1, 0, 0, 1};
        if (steps == 0) {
            divmod(c, d, q, r);
            Poly n10 = add(M.m00, mul_fast(q, M.m10));
            Poly n11 = add(M.m01, mul_fast(q, M.m11));
            M.m00.swap(M.m10);
            M.m01.swap(M.m11);
            M.m10.swap(n10);
            M.m11.swap(n11);
            c.swap(d);
            d.swap(r);
            continue;
        }
        lin2(W.m00, c, W.m01, d, t0);
        lin2(W.m10, c, W.m11, d, t1);
        c.swap(t0);
        d.swap(t1);
        lin2(W.m00, M.m00, W.m01, M.m10, t0);
        lin2(W.m00, M.m01, W.m01, M.m11, t1);
        lin2(W.m10, M.m00, W.m11, M.m10, t2);
        lin2(W.m10, M.m01, W.m11, M.m11, t3);
        M.m00.swap(t0);
        M.m01.swap(t1);
        M.m10.swap(t2);
        M.m11.swap(t3);
    }
}

}  // namespace

void xor_shift_words(u64 *a, const u64 *b, size_t bw, u64 sh) {
    size_t ws = (size_t)(sh >> 6);
    unsigned bs = (unsigned)(sh & 63);
    if (!bs) {
        for (size_t j = 0; j < bw; j++) a[j + ws] ^= b[j];
        return;
    }
    u64 carry = 0;
    for (size_t j = 0; j < bw; j++) {
        u64 v = b[j];
        a[j + ws] ^= (v << bs) | carry;
        carry = v >> (64 - bs);
    }
    a[bw + ws] ^= carry;
}

int host_threads() {
    static const int n = [] {
        int k = (int)std::thread::hardware_concurrency();
        cpu_set_t cs;
        if (sched_getaffinity(0, sizeof cs, &cs) == 0 && CPU_COUNT(&cs) > 0)
            k = std::min(k > 0 ? k : CPU_COUNT(&cs), CPU_COUNT(&cs));
        for (const char *v : {"GF2T_HOST_THREADS", "OMP_NUM_THREADS"}) {
            const char *e = getenv(v);
            if (e && atoi(e) > 0) {
                k = std::min(k > 0 ? k : atoi(e), atoi(e));
                break;
            }
        }
        return std::max(1, k);
    }();
    return n;
}

// deep-scan block size per r range: g_block_m[k] for 2^k <= r < 2^(k+1)
// (0 = the built-in choice)
static std::atomic<int> g_block_m[64];
int tuned_block_m(u64 r) {
    int k = 63 - __builtin_clzll(r | 1);
    return g_block_m[k].load(std::memory_order_relaxed);
}
bool set_tuned_block_m(int k, int m) {
    if (k < 2 || k > 40 || m < 0 || m > 64) return false;
    g_block_m[k].store(m);
    return true;
}

bool set_gcd_base_bits(int64_t bits) {
    if (bits < 128 || bits > 65536) return false;
    g_hgcd_base.store(bits);
    return true;
}
int64_t gcd_base_bits() { return g_hgcd_base.load(); }
bool set_school_words(int64_t words) {
    if (words < 8 || words > 512) return false;
    g_kbase.store((size_t)words);
    return true;
}
int64_t school_words() { return (int64_t)g_kbase.load(); }

void set_device_product_bits(u64 bits) { g_dev_bits.store(bits); }
u64 device_product_bits() { return g_dev_bits.load(); }

Poly mul_fast(const Poly &a, const Poly &b) {
    size_t na = words(a), nb = words(b);
    if (!na || !nb) return Poly();
    u64 lim = g_dev_bits.load(std::memory_order_relaxed);
    int64_t da = deg(a), db = deg(b);
    if (lim && da + 1 >= (int64_t)lim && db + 1 >= (int64_t)lim) {
        Poly out(gf2t_mul_out_words((u64)da + 1, (u64)db + 1), 0);
        PhaseTimer pt(t_stats.dev_calls, t_stats.dev_ns);
        check_rc(gf2t_mul_host(a.data(), (u64)da + 1, b.data(), (u64)db + 1, out.data()),
                 "gf2t_mul_host");
        trim(out);
        return out;
    }
    return host_mul(a.data(), na, b.data(), nb);
}

void hgcd(const Poly &a, const Poly &b, HMat &M, Poly &c, Poly &d, bool want_M) {
    int64_t n = deg(a);
    int64_t m = (n + 1) / 2;
    if (deg(b) < m) {
        M = HMat{Poly{1}, Poly(), Poly(), Poly{1}};
        c = a;
        d = b;
        trim(c);
        trim(d);
        return;
    }
    if (n < g_hgcd_base.load(std::memory_order_relaxed)) {
        PhaseTimer pt(t_stats.base_calls, t_stats.base_ns);
        hgcd_base(a, b, m, M, c, d);
        return;
    }
    // first half: the top n - m + 1 coefficients
    HMat R;
    Poly c0, d0;
    hgcd(shr(a, (u64)m), shr(b, (u64)m), R, c0, d0);
    Poly a1, b1;
    apply(R, low(a, (u64)m), low(b, (u64)m), a1, b1);
    xor_shift(a1, c0, (u64)m);
    xor_shift(b1, d0, (u64)m);
    trim(a1);
    trim(b1);
    if (deg(b1) < m) {
        M = std::move(R);
        c = std::move(a1);
        d = std::move(b1);
        return;
    }
    // one division step: (a1, b1) -> (b1, a1 mod b1)
    Poly q, rem;
    divmod(a1, b1, q, rem);
    HMat R2{R.m10, R.m11, add(R.m00, mul_fast(q, R.m10)), add(R.m01, mul_fast(q, R.m11))};
    if (deg(rem) < m) {
        M = std::move(R2);
        c = std::move(b1);
        d = std::move(rem);
        return;
    }
    // second half on the top 2(l - m) + 1 coefficients of (b1, rem)
    int64_t l = deg(b1), k = 2 * m - l;
    HMat S;
    Poly c3, d3;
    hgcd(shr(b1, (u64)k), shr(rem, (u64)k), S, c3, d3);
    // S applied to the low parts of (b1, rem) and, for M = S * R2, to the
    // columns (R2.m00, R2.m10) and (R2.m01, R2.m11) of R2, with one set of
    // transforms of S
    const Poly lb = low(b1, (u64)k), lr = low(rem, (u64)k);
    if (want_M) {
        const Poly *xs[3] = {&lb, &R2.m00, &R2.m01}, *ys[3] = {&lr, &R2.m10, &R2.m11};
        Poly *oxs[3] = {&c, &M.m00, &M.m01}, *oys[3] = {&d, &M.m10, &M.m11};
        applyN(S, 3, xs, ys, oxs, oys);
    } else {
        apply(S, lb, lr, c, d);
    }
    xor_shift(c, c3, (u64)k);
    xor_shift(d, d3, (u64)k);
    trim(c);
    trim(d);
}

Poly gcd_fast(const Poly &a0, const Poly &b0) {
    struct Caller {   // this thread computes: helpers are budgeted around it
        Caller() { g_gcd_callers.fetch_add(1); }
        ~Caller() { g_gcd_callers.fetch_sub(1); }
    } caller;
    struct Report {   // GF2T_GCD_STATS: one line per large gcd
        long long t0;
        int64_t bits;
        Report(int64_t b) : t0(gcd_stats_on() ? now_ns() : 0), bits(b) { t_stats = GcdStats{}; }
        ~Report() {
            if (!t0 || bits < (1 << 20)) return;
            const GcdStats &S = t_stats;
            fprintf(stderr, "gcd stats: %lld bits, %.3f s; device product calls %lld, %.3f s; "
                    "base case %lld calls, %.3f s\n", (long long)bits, (now_ns() - t0) * 1e-9,
                    S.dev_calls, S.dev_ns * 1e-9, S.base_calls, S.base_ns * 1e-9);
        }
    } report(std::max(deg(a0), deg(b0)) + 1);
    Poly a = a0, b = b0;
    trim(a);
    trim(b);
    if (deg(a) < deg(b)) std::swap(a, b);
    if (deg(b) >= 0 && deg(a) == deg(b)) {   // make deg a > deg b
        Poly t = add(a, b);
        a.swap(b);
        b.swap(t);
        if (deg(a) < deg(b)) std::swap(a, b);
    }
    while (deg(b) >= GCD_BASE) {
        HMat M;
        Poly c, d;
        hgcd(a, b, M, c, d, false);   // only the remainders are needed here
        if (deg(d) < 0) return c;
        Poly q, r;
        divmod(c, d, q, r);
        a.swap(d);
        b.swap(r);
    }
    return gcd(a, b);
}

}  // namespace gf2t_host
