// scan.cpp -- deep scan of sieve survivors: for each s, the least factor
// degree d in (k0, rhigh] of x^r + x^s + 1 and its numerically least
// degree-d factor, or 'u' / 'primitive' when the bound is reached
// (reference semantics: src/trinomial_scan.cu:240-381,
// src/tsfactor.cu:1089-1348; result lines src/tsfactor.cu:2480-2509).
//
// MI355X form: the s values are scanned in cohorts that advance in
// lockstep, one degree at a time, as slabs on the device:
//     F_k = F_{k-1}^2 mod T_s              (gf2t_square_fold_slab)
//     ACC = ACC * (F_k + x) mod T_s        (gf2t_mulmod_slab)
// for every live s at once.  At each interval end the ACC slab comes to
// the host, where a thread pool takes gcd(ACC, T_s): every irreducible
// factor with degree dividing some k of the interval divides it, and
// since all earlier intervals were clear, its factors are exactly those
// with degree in the interval.  A hit is resolved on the host from the
// small interval gcd: one factor (deg g < 2 ka) is the answer; otherwise
// the least degree comes from x^(2^k) mod g and the least mask from an
// equal-degree split.  Finished s values leave the cohort (the slab is
// compacted), so device work follows the survivors.
#include "../../include/gf2t.h"
#include "host_poly.hpp"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using namespace gf2t_host;

namespace gf2t {
void set_error(const char *fmt, ...);
}

struct gf2t_scan_results {
    std::vector<gf2t_scan_result> rows;
    std::vector<std::string> hex;
};

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// runs fn(0..n-1) on up to `threads` threads; the first exception a
// call throws stops the remaining work and is rethrown here
void parallel_for(size_t n, int threads, const std::function<void(size_t)> &fn) {
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
    ~DevMem() { if (p) gf2t_free(p); }
    int alloc(size_t bytes) {
        if (p) gf2t_free(p);
        p = nullptr;
        return gf2t_malloc(&p, bytes ? bytes : 8);
    }
    uint64_t *w() { return (uint64_t *)p; }
};

#define SCAN_CHECK(call)                                                       \
    do {                                                                       \
        int rc__ = (call);                                                     \
        if (rc__) return rc__;                                                 \
    } while (0)

struct Live {
    size_t idx;       // index into the caller's s list
    u64 s;
};

int run_cohort(u64 r, const gf2t_scan_params &P, u64 rhigh, bool limited,
               std::vector<Live> live, std::vector<gf2t_scan_result> &rows,
               std::vector<std::string> &hexes) {
    const size_t nw = (size_t)((r + 63) >> 6);
    const size_t pitch = (nw + 31) & ~(size_t)31;
    const int B = (int)live.size();
    DevMem F0, F1, ACC, SNAP, SV, WORK;
    SCAN_CHECK(F0.alloc(pitch * B * 8));
    SCAN_CHECK(F1.alloc(pitch * B * 8));
    SCAN_CHECK(ACC.alloc(pitch * B * 8));
    SCAN_CHECK(SNAP.alloc(pitch * B * 8));
    SCAN_CHECK(SV.alloc(B * 8));
    SCAN_CHECK(WORK.alloc(gf2t_mulmod_slab_workspace_bytes(r, B)));
    std::vector<u64> host((size_t)B * pitch, 0), one((size_t)B * pitch, 0);
    for (int i = 0; i < B; i++) {
        host[(size_t)i * pitch] = 2;   // F = x
        one[(size_t)i * pitch] = 1;    // ACC = 1
    }
    SCAN_CHECK(gf2t_memcpy_h2d(F0.p, host.data(), host.size() * 8, nullptr));
    std::vector<u64> svals(B);
    for (int i = 0; i < B; i++) svals[i] = live[i].s;
    SCAN_CHECK(gf2t_memcpy_h2d(SV.p, svals.data(), (size_t)B * 8, nullptr));
    uint64_t *cur = F0.w(), *oth = F1.w();
    for (u64 k = 0; k < P.k0; k++) {
        SCAN_CHECK(gf2t_square_fold_slab(cur, oth, pitch, r, SV.w(), B, nullptr));
        std::swap(cur, oth);
    }
    int n = B;
    u64 k = P.k0, L = P.L0 ? P.L0 : 64;
    const u64 Lmax = P.Lmax ? P.Lmax : 4096;
    std::vector<u64> acc_host;
    while (n > 0) {
        u64 ka = k + 1;
        if (ka > rhigh) break;
        u64 kb = std::min(k + L, rhigh);
        SCAN_CHECK(gf2t_memcpy_d2d(SNAP.p, cur, (size_t)n * pitch * 8, nullptr));
        SCAN_CHECK(gf2t_memcpy_h2d(ACC.p, one.data(), (size_t)n * pitch * 8, nullptr));
        for (u64 deg_k = ka; deg_k <= kb; deg_k++) {
            SCAN_CHECK(gf2t_square_fold_slab(cur, oth, pitch, r, SV.w(), n, nullptr));
            std::swap(cur, oth);
            SCAN_CHECK(gf2t_mulmod_slab(ACC.w(), pitch, cur, pitch, 1, r, SV.w(), ACC.w(), pitch,
                                        n, WORK.p, nullptr));
        }
        acc_host.assign((size_t)n * pitch, 0);
        SCAN_CHECK(gf2t_memcpy_d2h(acc_host.data(), ACC.p, acc_host.size() * 8, nullptr));
        std::vector<Poly> g(n);
        parallel_for((size_t)n, P.gcd_threads > 0 ? P.gcd_threads : 8, [&](size_t i) {
            Poly a = from_words(acc_host.data() + i * pitch, nw);
            g[i] = gcd_fast(trinomial(r, live[i].s), a);
        });
        // resolve hits (all earlier intervals of these s were clear)
        std::vector<int> hit;
        for (int i = 0; i < n; i++)
            if (deg(g[i]) > 0) hit.push_back(i);
        std::vector<u64> snap_host;
        if (!hit.empty()) {
            snap_host.assign((size_t)n * pitch, 0);
            SCAN_CHECK(gf2t_memcpy_d2h(snap_host.data(), SNAP.p, snap_host.size() * 8, nullptr));
        }
        std::vector<char> done(n, 0);
        parallel_for(hit.size(), P.gcd_threads > 0 ? P.gcd_threads : 8, [&](size_t h) {
            int i = hit[h];
            gf2t_scan_result &R = rows[live[i].idx];
            const Poly &gi = g[i];
            u64 gd = (u64)deg(gi);
            u64 d = 0;
            Poly mask;
            bool ok;
            if (gd < 2 * ka) {                       // a single irreducible factor
                d = gd;
                mask = gi;
                ok = true;
            } else {
                Poly h0 = from_words(snap_host.data() + (size_t)i * pitch, nw);
                ok = resolve_least(gi, ka, kb, h0, d, mask);
            }
            if (ok) {
                R.kind = GF2T_SCAN_FOUND;
                R.d = d;
                hexes[live[i].idx] = hex(mask);
            } else {
                R.kind = GF2T_SCAN_RANGE;
                R.rlo = ka;
                R.rhi = kb;
            }
            R.gcd_degree = gd;
            done[i] = 1;
        });
        if (kb >= rhigh)
            for (int i = 0; i < n; i++)
                if (!done[i]) {
                    gf2t_scan_result &R = rows[live[i].idx];
                    R.kind = limited ? GF2T_SCAN_GAVE_UP : GF2T_SCAN_PRIMITIVE;
                    done[i] = 1;
                }
        // compact the survivors of this interval to the front of the slabs
        int m = 0;
        for (int i = 0; i < n; i++) {
            if (done[i]) continue;
            if (m != i) {
                SCAN_CHECK(gf2t_memcpy_d2d(cur + (size_t)m * pitch, cur + (size_t)i * pitch,
                                           nw * 8, nullptr));
                live[m] = live[i];
                svals[m] = svals[i];
            }
            m++;
        }
        if (m != n && m > 0)
            SCAN_CHECK(gf2t_memcpy_h2d(SV.p, svals.data(), (size_t)m * 8, nullptr));
        if (P.verbose)
            fprintf(stderr, "  interval (%llu,%llu]: %d live, %zu hits, %d continue\n",
                    (unsigned long long)k, (unsigned long long)kb, n, hit.size(), m);
        n = m;
        k = kb;
        L = std::min(2 * L, Lmax);
    }
    return GF2T_OK;
}

}  // namespace

extern "C" int gf2t_scan_run(uint64_t r, const uint64_t *svals, size_t n,
                             const gf2t_scan_params *params, gf2t_scan_results **out) {
    if (!out || (n && !svals) || !params || r < 5) {
        gf2t::set_error("gf2t_scan_run: bad arguments");
        return GF2T_EINVAL;
    }
    *out = nullptr;
    gf2t_scan_params P = *params;
    if (P.k0 < 1) {
        gf2t::set_error("gf2t_scan_run: k0 (the depth the sieve reached) must be >= 1");
        return GF2T_EINVAL;
    }
    if (r >= ((uint64_t)1 << 31)) {
        gf2t::set_error("gf2t_scan_run: r=%llu beyond the product size limit",
                        (unsigned long long)r);
        return GF2T_ERANGE;
    }
    auto *res = new gf2t_scan_results();
    res->rows.resize(n);
    res->hex.resize(n);
    // group by degree bound (Swan or maxd) so a cohort's intervals clamp
    // exactly at its bound
    struct Group { u64 rhigh; bool limited; std::vector<Live> live; };
    std::vector<Group> groups;
    for (size_t i = 0; i < n; i++) {
        u64 s = svals[i];
        gf2t_scan_result &R = res->rows[i];
        memset(&R, 0, sizeof R);
        R.s = s;
        if (!(s >= 1 && 2 * s < r)) {
            delete res;
            gf2t::set_error("gf2t_scan_run: need 1 <= s <= (r-1)/2 (s=%llu)", (unsigned long long)s);
            return GF2T_EINVAL;
        }
        bool swan;
        u64 rhigh = swan_rhigh(r, s, &swan);
        bool limited = false;
        if (P.maxd && P.maxd < rhigh) { rhigh = P.maxd; limited = true; }
        auto it = std::find_if(groups.begin(), groups.end(), [&](const Group &g) {
            return g.rhigh == rhigh && g.limited == limited;
        });
        if (it == groups.end()) { groups.push_back(Group{rhigh, limited, {}}); it = groups.end() - 1; }
        it->live.push_back(Live{i, s});
    }
    size_t nw = (size_t)((r + 63) >> 6);
    size_t per = 4 * ((nw + 31) & ~(size_t)31) * 8 + gf2t_mulmod_slab_workspace_bytes(r, 1);
    size_t freeb = 0, totb = 0;
    int rc = gf2t_mem_info(&freeb, &totb);
    if (rc) { delete res; return rc; }
    size_t cap = std::max<size_t>(1, std::min<size_t>(65535, (freeb / 2) / per));
    if (P.batch > 0) cap = std::min<size_t>(cap, (size_t)P.batch);
    for (auto &G : groups)
        for (size_t off = 0; off < G.live.size(); off += cap) {
            std::vector<Live> part(G.live.begin() + off,
                                   G.live.begin() + std::min(G.live.size(), off + cap));
            try {
                rc = run_cohort(r, P, G.rhigh, G.limited, part, res->rows, res->hex);
            } catch (const std::exception &e) {
                gf2t::set_error("gf2t_scan_run: %s", e.what());
                rc = GF2T_EHIP;
            }
            if (rc) { delete res; return rc; }
        }
    for (size_t i = 0; i < n; i++) res->rows[i].hex_len = res->hex[i].size();
    *out = res;
    return GF2T_OK;
}

extern "C" size_t gf2t_scan_results_count(const gf2t_scan_results *res) {
    return res ? res->rows.size() : 0;
}

extern "C" int gf2t_scan_results_get(const gf2t_scan_results *res, size_t i,
                                     gf2t_scan_result *row, char *hexbuf, size_t cap) {
    if (!res || i >= res->rows.size()) return GF2T_EINVAL;
    if (row) *row = res->rows[i];
    if (hexbuf) {
        const std::string &h = res->hex[i];
        if (cap < h.size() + 1) return GF2T_ERANGE;
        memcpy(hexbuf, h.c_str(), h.size() + 1);
    }
    return GF2T_OK;
}

extern "C" int gf2t_scan_line(const gf2t_scan_results *res, size_t i, char *buf, size_t cap) {
    if (!res || i >= res->rows.size() || !buf) return GF2T_EINVAL;
    const gf2t_scan_result &R = res->rows[i];
    std::string line = std::to_string(R.s);
    switch (R.kind) {
    case GF2T_SCAN_FOUND: line += " " + std::to_string(R.d) + " p" + res->hex[i]; break;
    case GF2T_SCAN_GAVE_UP: line += " u"; break;
    case GF2T_SCAN_PRIMITIVE: line += " primitive"; break;
    case GF2T_SCAN_RANGE:
        line += " r" + std::to_string(R.rlo) + "-" + std::to_string(R.rhi);
        break;
    default: line += " ?"; break;
    }
    if (cap < line.size() + 1) return GF2T_ERANGE;
    memcpy(buf, line.c_str(), line.size() + 1);
    return GF2T_OK;
}

extern "C" void gf2t_scan_results_free(gf2t_scan_results *res) { delete res; }

extern "C" int64_t gf2t_gcd_host(const uint64_t *a, size_t aw, const uint64_t *b, size_t bw,
                                 uint64_t *g, size_t gcap) {
    Poly G;
    try {
        G = gcd_fast(from_words(a, aw), from_words(b, bw));
    } catch (const std::exception &e) {
        gf2t::set_error("gf2t_gcd_host: %s", e.what());
        return GF2T_EHIP;
    }
    int64_t d = deg(G);
    if (g && gcap) {
        memset(g, 0, gcap * 8);
        memcpy(g, G.data(), std::min(gcap, G.size()) * 8);
    }
    return d;
}

extern "C" int gf2t_edf_least(const uint64_t *t, size_t tw, uint64_t d, uint64_t *out,
                              size_t cap) {
    Poly f = edf_least(from_words(t, tw), d);
    if (f.empty()) {
        gf2t::set_error("gf2t_edf_least: no degree-%llu factor found", (unsigned long long)d);
        return GF2T_EINTERNAL;
    }
    if (cap < f.size()) return GF2T_ERANGE;
    memset(out, 0, cap * 8);
    memcpy(out, f.data(), f.size() * 8);
    return GF2T_OK;
}

extern "C" uint64_t gf2t_swan_bound(uint64_t r, uint64_t s, int *swan) {
    bool sw = false;
    u64 b = swan_rhigh(r, s, &sw);
    if (swan) *swan = sw ? 1 : 0;
    return b;
}
