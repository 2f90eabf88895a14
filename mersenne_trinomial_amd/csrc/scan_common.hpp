// scan_common.hpp -- host pieces shared by the deep scan (scan.cpp) and the
// serving engine (serve.cpp): a small parallel-for, device allocations,
// and the resolution of an interval gcd into a certificate.
#pragma once
#include "../../include/gf2t.h"
#include "host_poly.hpp"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace gf2t {
void set_error(const char *fmt, ...);
}

namespace gf2t_scan {

using gf2t_host::Poly;
using gf2t_host::u64;

inline double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// runs fn(0..n-1) on up to `threads` threads; the first exception a
// call throws stops the remaining work and is rethrown here
inline void parallel_for(size_t n, int threads, const std::function<void(size_t)> &fn) {
    if (n == 0) return;
    int t = std::max(1, std::min<int>(threads, (int)n));
    std::atomic<size_t> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr err;
    std::mutex err_mu;
    std::vector<std::thread> pool;
    for (int i = 0; i < t; i++)
        pool.emplace_back([&] {
            for (;;) {
                size_t k = next.fetch_add(1);
                if (k >= n || failed.load()) break;
                try {
                    fn(k);
                } catch (...) {
                    std::lock_guard<std::mutex> lk(err_mu);
                    if (!err) err = std::current_exception();
                    failed.store(true);
                }
            }
        });
    for (auto &th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

struct DevMem {
    void *p = nullptr;
    DevMem() {}
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem() { if (p) gf2t_free(p); }
    int alloc(size_t bytes) {
        if (p) gf2t_free(p);
        p = nullptr;
        return gf2t_malloc(&p, bytes ? bytes : 8);
    }
    uint64_t *w() { return (uint64_t *)p; }
};

// x^(2^k) mod g on the host (g small: the interval gcd)
inline Poly frob_mod(const Poly &g, u64 k) {
    gf2t_host::Modulus G(g);
    Poly h = G.reduce(Poly{2});
    for (u64 i = 0; i < k; i++) h = G.sqr(h);
    return h;
}

// Outcome of an interval whose gcd g with T is nontrivial.  Every earlier
// interval was clear, so each irreducible factor of g has its degree in
// [ka, kb].  Returns false when the least degree lies beyond rhigh (a
// factor found past the bound is not reported).
// A factor of degree beyond rhigh (found in the last whole block) is
// reported only when past_bound is set; the range of an unresolved hit
// ends at rhigh, or at kb with past_bound.
inline bool resolve_hit(const Poly &g, u64 ka, u64 kb, u64 rhigh, gf2t_scan_result &R,
                        std::string &hex, bool past_bound = false) {
    const u64 gd = (u64)gf2t_host::deg(g);
    u64 d = 0;
    Poly mask;
    bool ok;
    if (gd < 2 * ka) {   // a single irreducible factor
        d = gd;
        mask = g;
        ok = true;
    } else {
        ok = gf2t_host::resolve_least(g, ka, kb, frob_mod(g, ka - 1), d, mask);
    }
    if (ok && d > rhigh && !past_bound) return false;
    if (ok) {
        R.kind = GF2T_SCAN_FOUND;
        R.d = d;
        hex = gf2t_host::hex(mask);
    } else {
        R.kind = GF2T_SCAN_RANGE;
        R.rlo = ka;
        R.rhi = past_bound ? kb : std::min(kb, rhigh);
    }
    R.gcd_degree = gd;
    return true;
}

// Advances a slab of n residues (residue y uses ds[y / group]) by `steps`
// squarings, ping-ponging between cur and oth (cur ends holding the
// result).  Two squarings per launch when the residues are large enough
// for the two-squaring tiles to be full (r >= 2^18): in work units when a
// plan with the groups' layouts is given, else in plain tiles.
inline int slab_square_steps(u64 *&cur, u64 *&oth, size_t pitch, u64 r, const u64 *ds, size_t n,
                             int group, u64 steps, void *stream,
                             const gf2t_sq2_slab_plan *plan = nullptr) {
    const bool two = r >= (1u << 18) && gf2t_host::square_fuse_enabled();
    for (u64 t = 0; t < steps;) {
        int rc;
        if (two && steps - t >= 2) {
            rc = plan ? gf2t_square2_fold_slab_p(plan, cur, oth, pitch, n, group, stream)
                      : gf2t_square2_fold_slab_g(cur, oth, pitch, r, ds, n, group, stream);
            t += 2;
        } else {
            rc = gf2t_square_fold_slab_g(cur, oth, pitch, r, ds, n, group, stream);
            t += 1;
        }
        if (rc) return rc;
        std::swap(cur, oth);
    }
    return GF2T_OK;
}

// degrees per product in the blocked scheme.  Per degree the scheme costs
// m slab squarings plus 1/m of a slab product, so the best m is about
// sqrt(product time / squaring time): ~16 for r near 2^17, ~32 from
// r ~ 2^19 on (ratios of about 300 and 1100, from the squaring rate of
// 4.7 TB/s and slab product times of ~5 us and ~8 ms per residue).
inline int block_for(int requested, u64 r) {
    int m = requested > 0 ? requested : gf2t_host::tuned_block_m(r);
    if (m <= 0) m = r < (1u << 14) ? 4 : r < (1u << 19) ? 16 : 32;
    if (m > 64) m = 64;
    while (m > 1 && 2 * (u64)m >= r) m /= 2;
    return m;
}

}  // namespace gf2t_scan
