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
#include "scan_common.hpp"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
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
using namespace gf2t_scan;


struct gf2t_scan_results {
    std::vector<gf2t_scan_result> rows;
    std::vector<std::string> hex;
    // wall seconds of the run's phases, summed over its cohorts: waiting
    // for the device, interval gcds, resolution of hits; and the number
    // of interval gcds taken
    double t_dev = 0, t_gcd = 0, t_res = 0;
    uint64_t n_gcd = 0;
};

namespace {

// GF2T_SCAN_GRAPHS=0 issues every block's launches directly (for timing
// the graphs against plain launches); default on
bool scan_graphs_enabled() {
    static const bool on = [] {
        const char *e = getenv("GF2T_SCAN_GRAPHS");
        return !(e && *e == '0');
    }();
    return on;
}

// Staggered sub-cohorts: one cohort's s values split into parts that run
// on their own threads and streams, part j starting its device work when
// part j-1 has the products of its first interval.  The host then takes
// the gcds of one part while the device runs the next, instead of every
// gcd waiting for the whole cohort's device work.  `mu` guards the
// results' shared phase counters.
struct Stagger {
    std::mutex mu;
    std::condition_variable cv;
    int next = 0;   // the part whose turn it is to start
    void wait(int turn) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return next >= turn; });
    }
    void pass(int turn) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (next == turn) next = turn + 1;
        }
        cv.notify_all();
    }
};

std::atomic<int> g_scan_stagger{0};   // parts per cohort; 0 = automatic

#define SCAN_CHECK(call)                                                       \
    do {                                                                       \
        int rc__ = (call);                                                     \
        if (rc__) return rc__;                                                 \
    } while (0)

struct Live {
    size_t idx;       // index into the caller's s list
    u64 s;
};

// One cohort of s values sharing a degree bound rhigh, advanced together.
// block_m == 1: per degree one squaring of F and one product ACC *= F + x
// (the reference's trinomial_scan loop, src/trinomial_scan.cu:345-380);
// block_m > 1: the blocked scheme of ddf.hip (tsfactor's h-table,
// src/tsfactor.cu:1242-1341), one product per m degrees.  Intervals end
// with a host gcd(ACC, T_s) per residue.
// sg (may be null): this cohort is part `turn` of a staggered cohort
int run_cohort(u64 r, const gf2t_scan_params &P, u64 rhigh, bool limited,
               std::vector<Live> live, gf2t_scan_results &out, bool &stopped,
               Stagger *sg = nullptr, int turn = 0) {
    std::vector<gf2t_scan_result> &rows = out.rows;
    std::vector<std::string> &hexes = out.hex;
    // the next part's turn comes once this one has its first interval's
    // products, or when it ends (an error, nothing to scan)
    struct TurnGuard {
        Stagger *sg;
        int turn;
        bool passed = false;
        void pass() {
            if (sg && !passed) sg->pass(turn);
            passed = true;
        }
        ~TurnGuard() { pass(); }
    } turn_guard{sg, turn};
    if (P.k0 >= rhigh) {   // nothing left to scan: the bound is at or below the sieve depth
        for (const Live &L : live)
            rows[L.idx].kind = limited ? GF2T_SCAN_GAVE_UP : GF2T_SCAN_PRIMITIVE;
        return GF2T_OK;
    }
    const size_t nw = (size_t)((r + 63) >> 6);
    const size_t pitch = (nw + 31) & ~(size_t)31;
    const int B = (int)live.size();
    const int m = block_for(P.block_m, r);
    const size_t tab = (m > 1 ? (size_t)m : 1) * pitch;   // words per residue in F / h-table
    DevMem T0, T1, ACC, C, SV, WORK;
    SCAN_CHECK(T0.alloc(tab * B * 8));
    SCAN_CHECK(T1.alloc(tab * B * 8));
    SCAN_CHECK(ACC.alloc(pitch * B * 8));
    if (m > 1) SCAN_CHECK(C.alloc(pitch * B * 8));
    SCAN_CHECK(SV.alloc(B * 8));
    SCAN_CHECK(WORK.alloc(gf2t_mulmod_slab_workspace_bytes(r, B)));
    // the cohort's own stream: the device work of a block can then be
    // captured once as a HIP graph and replayed (see issue_block)
    void *st = nullptr;
    SCAN_CHECK(gf2t_stream_create(&st));
    struct StreamGuard {
        void *s;
        ~StreamGuard() {
            gf2t_stream_sync(s);
            gf2t_stream_destroy(s);
        }
    } stream_guard{st};
    std::vector<u64> svals(B);
    for (int i = 0; i < B; i++) svals[i] = live[i].s;
    SCAN_CHECK(gf2t_memcpy_h2d(SV.p, svals.data(), (size_t)B * 8, st));
    // the slots' tile layouts for the two-squaring work units (r >= 2^18)
    gf2t_sq2_slab_plan *sq2 = nullptr;
    if (r >= (1u << 18) && B > 0)
        SCAN_CHECK(gf2t_sq2_slab_plan_create(r, svals.data(), (size_t)B, &sq2));
    struct PlanGuard {   // destroyed after the stream is synchronized (declared before it)
        gf2t_sq2_slab_plan *p;
        ~PlanGuard() { gf2t_sq2_slab_plan_destroy(p); }
    } plan_guard{sq2};
    uint64_t *cur = T0.w(), *oth = T1.w();
    if (sg) sg->wait(turn);   // staggered: the previous part has its first products
    double t_dev = 0, t_gcd = 0, t_res = 0, t0 = now_s();
    uint64_t n_gcd = 0;
    struct AddTimes {   // the cohort's phase times go into the results on every exit
        gf2t_scan_results &o;
        double &a, &b, &c;
        uint64_t &g;
        Stagger *sg;
        ~AddTimes() {
            std::unique_lock<std::mutex> lk;
            if (sg) lk = std::unique_lock<std::mutex>(sg->mu);
            o.t_dev += a; o.t_gcd += b; o.t_res += c; o.n_gcd += g;
        }
    } add_times{out, t_dev, t_gcd, t_res, n_gcd, sg};
    if (m > 1) {
        // h-table for the first block, degrees k0+1 .. k0+m
        SCAN_CHECK(gf2t_ddf_init_slab(cur, oth, pitch, m, r, SV.w(), P.k0 + 1, B, st));
    } else {
        std::vector<u64> host((size_t)B * pitch, 0);
        for (int i = 0; i < B; i++) host[(size_t)i * pitch] = 2;   // F = x
        SCAN_CHECK(gf2t_memcpy_h2d(cur, host.data(), host.size() * 8, st));
        SCAN_CHECK(slab_square_steps(cur, oth, pitch, r, SV.w(), (size_t)B, 1, P.k0, st, sq2));
    }
    // One block of the interval on the first nslot slots: m degrees of the
    // blocked scheme (Horner product C, ACC *= C, m squarings of the
    // h-table), or one degree of the plain one (F squared, ACC *= F + x).
    // Its launches depend only on (cur, nslot), so after the first block
    // it is captured as a HIP graph per (cur, nslot) and replayed: one
    // graph launch instead of ~60 kernel launches per block.
    auto issue_block = [&](uint64_t *&c, uint64_t *&o, int nslot) -> int {
        if (m > 1) {
            SCAN_CHECK(gf2t_horner_slab(c, pitch, m, r, SV.w(), C.w(), pitch, nslot, st));
            SCAN_CHECK(gf2t_mulmod_slab(ACC.w(), pitch, C.w(), pitch, 0, r, SV.w(), ACC.w(), pitch,
                                        nslot, WORK.p, st));
            SCAN_CHECK(
                slab_square_steps(c, o, pitch, r, SV.w(), (size_t)nslot * m, m, (u64)m, st, sq2));
        } else {
            SCAN_CHECK(gf2t_square_fold_slab(c, o, pitch, r, SV.w(), nslot, st));
            std::swap(c, o);
            SCAN_CHECK(gf2t_mulmod_slab(ACC.w(), pitch, c, pitch, 1, r, SV.w(), ACC.w(), pitch,
                                        nslot, WORK.p, st));
        }
        return GF2T_OK;
    };
    struct BlockGraph {
        uint64_t *cin, *cout, *oout;
        int nslot;
        gf2t_graph *g;
    };
    std::vector<BlockGraph> graphs;
    struct GraphGuard {
        std::vector<BlockGraph> &v;
        ~GraphGuard() {
            for (auto &x : v) gf2t_graph_destroy(x.g);
        }
    } graph_guard{graphs};
    // a staggered part issues its launches directly: its sibling parts'
    // gcd threads allocate and synchronize while it would be capturing
    const bool use_graphs = scan_graphs_enabled() && !sg;
    bool warmed = false;
    auto run_block = [&](int nslot) -> int {
        if (!use_graphs || !warmed) {   // the first block runs eagerly (plans, constants)
            warmed = true;
            return issue_block(cur, oth, nslot);
        }
        for (auto it = graphs.begin(); it != graphs.end();)
            if (it->nslot != nslot) {   // the cohort was compacted: older graphs are stale
                gf2t_graph_destroy(it->g);
                it = graphs.erase(it);
            } else {
                ++it;
            }
        BlockGraph *G = nullptr;
        for (auto &x : graphs)
            if (x.cin == cur) G = &x;
        if (!G) {
            uint64_t *c = cur, *o = oth;
            SCAN_CHECK(gf2t_graph_begin(st));
            int rc = issue_block(c, o, nslot);
            gf2t_graph *g = nullptr;
            int rc2 = gf2t_graph_end(st, &g);
            if (rc) {
                gf2t_graph_destroy(g);
                return rc;
            }
            if (rc2) return rc2;
            graphs.push_back(BlockGraph{cur, c, o, nslot, g});
            G = &graphs.back();
        }
        SCAN_CHECK(gf2t_graph_launch(G->g, st));
        cur = G->cout;
        oth = G->oout;
        return GF2T_OK;
    };
    // Work still queued on the device when the cohort ends (a speculative
    // interval) must finish before the buffers above are freed; declared
    // after them, this runs first.
    struct SyncOnExit {
        void *st;
        ~SyncOnExit() { gf2t_stream_sync(st); }
    } sync_on_exit{st};
    int n = B;
    u64 k = P.k0, L = P.L0 ? P.L0 : 64;
    const u64 Lmax = P.Lmax ? P.Lmax : 4096;
    const int threads = P.gcd_threads > 0 ? P.gcd_threads : 8;
    std::vector<u64> acc_host;
    // the device work of the interval after frontier k0 (length L0, on the
    // first nslot slots): its products accumulate in ACC; returns its end
    auto enqueue = [&](u64 k0, u64 L0, int nslot, u64 &kb) -> int {
        const u64 ka = k0 + 1;
        const u64 kend = std::min(k0 + L0, rhigh);
        kb = kend;
        SCAN_CHECK(gf2t_set_one_slab(ACC.w(), pitch, nw, nslot, st));
        if (m > 1) {
            // whole blocks: the last may run past rhigh; factors found
            // beyond it are not reported
            u64 nblk = (kend - k0 + (u64)m - 1) / (u64)m;
            kb = k0 + nblk * (u64)m;
            for (u64 b = 0; b < nblk; b++) SCAN_CHECK(run_block(nslot));
        } else {
            for (u64 deg_k = ka; deg_k <= kb; deg_k++) SCAN_CHECK(run_block(nslot));
        }
        return GF2T_OK;
    };
    if (k + 1 > rhigh) return GF2T_OK;
    u64 kb = 0;
    SCAN_CHECK(enqueue(k, L, n, kb));
    while (n > 0) {
        // the interval (k, kb] of the n live slots is queued: wait for its
        // products
        acc_host.assign((size_t)n * pitch, 0);
        SCAN_CHECK(gf2t_memcpy_d2h(acc_host.data(), ACC.p, acc_host.size() * 8, st));
        double t1 = now_s();
        t_dev += t1 - t0;
        turn_guard.pass();
        // speculation (src/tsfactor.cu:1279-1331): the next interval goes
        // to the device for every slot while the host takes this one's
        // gcds; the slots that finish here simply drop its results.  An s
        // clear through degree k is still clear through kb with
        // probability about k / kb, so when that is below a quarter (the
        // first interval after the sieve depth) most of the speculative
        // work would be dropped, and it would slow the gcds' own device
        // products: the next interval is then queued after the gcds, for
        // the slots that continue.
        const u64 ka = k + 1, Ln = std::min(2 * L, Lmax);
        const bool more = kb < rhigh;
        const bool spec = more && 4 * k >= kb;
        u64 kb_next = 0;
        if (spec) SCAN_CHECK(enqueue(kb, Ln, n, kb_next));
        std::vector<Poly> g(n);
        n_gcd += (uint64_t)n;
        parallel_for((size_t)n, threads, [&](size_t i) {
            Poly a = from_words(acc_host.data() + i * pitch, nw);
            g[i] = gcd_fast(trinomial(r, live[i].s), a);
        });
        // resolve hits: every earlier interval of these s was clear, so
        // each irreducible factor of g has its degree in [ka, kb]
        std::vector<int> hit;
        for (int i = 0; i < n; i++)
            if (deg(g[i]) > 0) hit.push_back(i);
        double t2 = now_s();
        t_gcd += t2 - t1;
        std::vector<char> done(n, 0);
        parallel_for(hit.size(), threads, [&](size_t h) {
            int i = hit[h];
            if (resolve_hit(g[i], ka, kb, rhigh, rows[live[i].idx], hexes[live[i].idx],
                            P.past_bound != 0))
                done[i] = 1;
        });
        t0 = now_s();
        t_res += t0 - t2;
        if (!more)
            for (int i = 0; i < n; i++)
                if (!done[i]) {
                    gf2t_scan_result &R = rows[live[i].idx];
                    R.kind = limited ? GF2T_SCAN_GAVE_UP : GF2T_SCAN_PRIMITIVE;
                    done[i] = 1;
                }
        // compact the survivors to the front of the slabs: their advanced
        // Frobenius state and, when the next interval was queued
        // speculatively, its products (the copies queue behind that
        // interval's launches)
        int nl = 0;
        for (int i = 0; i < n; i++) {
            if (done[i]) continue;
            if (nl != i) {
                SCAN_CHECK(gf2t_memcpy_d2d(cur + (size_t)nl * tab, cur + (size_t)i * tab,
                                           tab * 8, st));
                if (spec)
                    SCAN_CHECK(gf2t_memcpy_d2d(ACC.w() + (size_t)nl * pitch,
                                               ACC.w() + (size_t)i * pitch, pitch * 8, st));
                live[nl] = live[i];
                svals[nl] = svals[i];
            }
            nl++;
        }
        if (nl != n && nl > 0) {
            SCAN_CHECK(gf2t_memcpy_h2d(SV.p, svals.data(), (size_t)nl * 8, st));
            if (sq2) SCAN_CHECK(gf2t_sq2_slab_plan_set(sq2, svals.data(), (size_t)nl, st));
        }
        if (P.verbose)
            fprintf(stderr, "  interval (%llu,%llu]: %d live, %zu hits, %d continue"
                    " (m=%d; device wait %.3fs, gcd %.3fs, resolve %.3fs so far)\n",
                    (unsigned long long)k, (unsigned long long)kb, n, hit.size(), nl, m, t_dev,
                    t_gcd, t_res);
        n = nl;
        if (more && !spec && n > 0) SCAN_CHECK(enqueue(kb, Ln, n, kb_next));
        k = kb;
        L = Ln;
        kb = kb_next;
        if (n > 0 && P.progress) {
            // every s still live is clear through degree k
            const u64 kdone = std::min(k, rhigh);
            if (P.progress(P.progress_user, kdone, svals.data(), (size_t)n)) {
                for (int i = 0; i < n; i++) rows[live[i].idx].kind = GF2T_SCAN_STOPPED;
                stopped = true;
                return GF2T_OK;
            }
        }
    }
    return GF2T_OK;
}

// Parts a cohort of n s values is staggered into: the "scan_stagger"
// knob when set; else, for residues of 2^22 bits and more (where one
// interval's device work per s is long against the start of a gcd), up
// to four; never with a progress callback (its calls and stop flag
// belong to one cohort at a time).
int stagger_parts(u64 r, size_t n, const gf2t_scan_params &P) {
    if (P.progress || n < 2) return 1;
    const int v = g_scan_stagger.load(std::memory_order_relaxed);
    if (v > 0) return (int)std::min<size_t>((size_t)v, n);
    if (v < 0 || r < ((u64)1 << 22)) return 1;
    return (int)std::min<size_t>(4, n);
}

// The s values of one cohort in K contiguous parts, each run by its own
// thread as a cohort of its own, part j starting its device work when
// part j-1 has its first interval's products (Stagger).
int run_staggered(u64 r, const gf2t_scan_params &P, u64 rhigh, bool limited,
                  const std::vector<Live> &live, int K, gf2t_scan_results &out) {
    Stagger sg;
    std::vector<int> rcs((size_t)K, GF2T_OK);
    std::vector<std::string> errs((size_t)K);
    std::vector<std::thread> th;
    const size_t n = live.size();
    for (int j = 0; j < K; j++) {
        const size_t a = n * (size_t)j / (size_t)K, b = n * (size_t)(j + 1) / (size_t)K;
        std::vector<Live> part(live.begin() + a, live.begin() + b);
        th.emplace_back([&, j, part]() mutable {
            bool stopped = false;
            try {
                rcs[j] = run_cohort(r, P, rhigh, limited, std::move(part), out, stopped, &sg, j);
                if (rcs[j]) errs[j] = gf2t_last_error();
            } catch (const std::exception &e) {
                rcs[j] = GF2T_EHIP;
                errs[j] = e.what();
            }
        });
    }
    for (auto &t : th) t.join();
    for (int j = 0; j < K; j++)
        if (rcs[j]) {
            gf2t::set_error("gf2t_scan_run: %s", errs[j].c_str());
            return rcs[j];
        }
    return GF2T_OK;
}

}  // namespace

namespace gf2t_host {
void set_scan_stagger(int parts) { g_scan_stagger.store(parts); }
int scan_stagger() { return g_scan_stagger.load(); }
}  // namespace gf2t_host

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
    const size_t pitch = (nw + 31) & ~(size_t)31;
    const size_t per = (2 * (size_t)block_for(P.block_m, r) + 2) * pitch * 8 +
                       gf2t_mulmod_slab_workspace_bytes(r, 1);
    size_t freeb = 0, totb = 0;
    int rc = gf2t_mem_info(&freeb, &totb);
    if (rc) { delete res; return rc; }
    size_t cap = std::max<size_t>(1, std::min<size_t>(65535, (freeb / 2) / per));
    if (P.batch > 0) cap = std::min<size_t>(cap, (size_t)P.batch);
    bool stopped = false;
    for (auto &G : groups)
        for (size_t off = 0; off < G.live.size(); off += cap) {
            std::vector<Live> part(G.live.begin() + off,
                                   G.live.begin() + std::min(G.live.size(), off + cap));
            if (stopped) {
                for (const Live &L : part) res->rows[L.idx].kind = GF2T_SCAN_STOPPED;
                continue;
            }
            const int K = stagger_parts(r, part.size(), P);
            if (K > 1) {
                rc = run_staggered(r, P, G.rhigh, G.limited, part, K, *res);
            } else {
                try {
                    rc = run_cohort(r, P, G.rhigh, G.limited, part, *res, stopped);
                } catch (const std::exception &e) {
                    gf2t::set_error("gf2t_scan_run: %s", e.what());
                    rc = GF2T_EHIP;
                }
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

extern "C" int gf2t_scan_results_times(const gf2t_scan_results *res, double *dev_s,
                                       double *gcd_s, double *resolve_s, uint64_t *n_gcd) {
    if (!res) return GF2T_EINVAL;
    if (dev_s) *dev_s = res->t_dev;
    if (gcd_s) *gcd_s = res->t_gcd;
    if (resolve_s) *resolve_s = res->t_res;
    if (n_gcd) *n_gcd = res->n_gcd;
    return GF2T_OK;
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
