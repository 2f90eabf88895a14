// serve.cpp -- deep-scan server with continuous batching: s values of many
// jobs (sieve survivors of one r) share one device slab and advance
// together in rounds of m degrees (the blocked scheme of ddf.hip), each
// with its own frontier.  New s values are admitted into free slots at
// any round boundary, finished ones leave at their own interval ends, and
// the slab's size is fixed from an HBM budget, so admission is by device
// footprint.  No reference counterpart: the reference scans one s at a
// time (src/trinomial_scan.cu:240-381, src/tsfactor.cu:1089-1348); each
// s gets the same certificate line as in gf2t_scan_run, alone or in any
// mix.
#include "../../include/gf2t.h"
#include "host_poly.hpp"
#include "scan_common.hpp"
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

using namespace gf2t_host;
using namespace gf2t_scan;

namespace {

struct Slot {
    u64 job, s, k0, rhigh;
    bool limited;
    u64 kdone;        // degrees folded into ACC so far
    u64 ka, kb;       // current interval
    u64 L;            // current interval length
};

struct Done {
    u64 job;
    std::string line;
};

#define SV_CHECK(call)                                                         \
    do {                                                                       \
        int rc__ = (call);                                                     \
        if (rc__) return rc__;                                                 \
    } while (0)

}  // namespace

struct gf2t_server {
    u64 r = 0;
    gf2t_scan_params P{};
    int m = 1;
    size_t nw = 0, pitch = 0, tab = 0, cap = 0, per = 0;
    DevMem T0, T1, ACC, C, SV, WORK;
    uint64_t *cur = nullptr, *oth = nullptr;
    std::vector<Slot> live;   // slot i of the slabs
    std::vector<u64> svals;   // s of slot i (mirrors SV)
    // the slots' two-squaring tile layouts (r >= 2^18), reset from svals
    // before a round when they changed
    gf2t_sq2_slab_plan *sq2 = nullptr;
    bool sq2_stale = true;
    std::deque<Slot> queue;
    std::deque<Done> done;
    std::mutex mu;            // one caller at a time

    int create(u64 r_, const gf2t_scan_params &p, size_t mem) {
        r = r_;
        P = p;
        m = block_for(P.block_m, r);
        if (m < 2) m = 2;
        nw = (size_t)((r + 63) >> 6);
        pitch = (nw + 31) & ~(size_t)31;
        tab = (size_t)m * pitch;
        per = (2 * tab + 2 * pitch) * 8 + gf2t_mulmod_slab_workspace_bytes(r, 1) + 8;
        if (!mem) {
            size_t fb = 0, tb = 0;
            SV_CHECK(gf2t_mem_info(&fb, &tb));
            mem = fb / 2;
        }
        cap = std::min<size_t>(65535, mem / per);
        if (cap < 1) {
            gf2t::set_error("gf2t_server_create: %zu bytes hold no slot of %zu bytes", mem, per);
            return GF2T_ENOMEM;
        }
        SV_CHECK(T0.alloc(tab * cap * 8));
        SV_CHECK(T1.alloc(tab * cap * 8));
        SV_CHECK(ACC.alloc(pitch * cap * 8));
        SV_CHECK(C.alloc(pitch * cap * 8));
        SV_CHECK(SV.alloc(cap * 8));
        SV_CHECK(WORK.alloc(gf2t_mulmod_slab_workspace_bytes(r, (int)cap)));
        cur = T0.w();
        oth = T1.w();
        if (r >= (1u << 18)) {
            const std::vector<u64> one(cap, 1);
            SV_CHECK(gf2t_sq2_slab_plan_create(r, one.data(), cap, &sq2));
        }
        return GF2T_OK;
    }
    ~gf2t_server() {
        (void)gf2t_stream_sync(nullptr);
        gf2t_sq2_slab_plan_destroy(sq2);
    }

    void finish(const Slot &q, const gf2t_scan_result &R, const std::string &hex) {
        std::string line = std::to_string(q.s);
        switch (R.kind) {
        case GF2T_SCAN_FOUND: line += " " + std::to_string(R.d) + " p" + hex; break;
        case GF2T_SCAN_GAVE_UP: line += " u"; break;
        case GF2T_SCAN_PRIMITIVE: line += " primitive"; break;
        default: line += " r" + std::to_string(R.rlo) + "-" + std::to_string(R.rhi); break;
        }
        done.push_back(Done{q.job, line});
    }

    void next_interval(Slot &q) {
        q.ka = q.kdone + 1;
        u64 kend = std::min(q.kdone + q.L, q.rhigh);
        q.kb = q.kdone + (kend - q.kdone + (u64)m - 1) / (u64)m * (u64)m;
    }

    int submit(u64 job, const u64 *s, size_t n, u64 k0, u64 maxd) {
        for (size_t i = 0; i < n; i++) {
            if (!(s[i] >= 1 && 2 * s[i] < r)) {
                gf2t::set_error("gf2t_server_submit: need 1 <= s <= (r-1)/2 (s=%llu)",
                                (unsigned long long)s[i]);
                return GF2T_EINVAL;
            }
        }
        for (size_t i = 0; i < n; i++) {
            bool swan;
            Slot q{};
            q.job = job;
            q.s = s[i];
            q.k0 = k0;
            q.rhigh = swan_rhigh(r, s[i], &swan);
            q.limited = false;
            if (maxd && maxd < q.rhigh) { q.rhigh = maxd; q.limited = true; }
            q.kdone = k0;
            q.L = P.L0 ? P.L0 : 64;
            if (q.kdone >= q.rhigh) {
                gf2t_scan_result R{};
                R.kind = q.limited ? GF2T_SCAN_GAVE_UP : GF2T_SCAN_PRIMITIVE;
                finish(q, R, std::string());
                continue;
            }
            next_interval(q);
            queue.push_back(q);
        }
        return GF2T_OK;
    }

    // fill free slots from the queue; each run of equal k0 gets one
    // h-table build
    int admit() {
        size_t n = live.size();
        size_t add = std::min(cap - n, queue.size());
        if (!add) return GF2T_OK;
        for (size_t i = 0; i < add; i++) {
            live.push_back(queue.front());
            svals.push_back(queue.front().s);
            queue.pop_front();
        }
        SV_CHECK(gf2t_memcpy_h2d(SV.w() + n, svals.data() + n, add * 8, nullptr));
        sq2_stale = true;
        size_t i = n;
        while (i < n + add) {
            size_t j = i;
            while (j < n + add && live[j].kdone == live[i].kdone) j++;
            SV_CHECK(gf2t_ddf_init_slab(cur + i * tab, oth + i * tab, pitch, m, r, SV.w() + i,
                                        live[i].kdone + 1, (int)(j - i), nullptr));
            i = j;
        }
        SV_CHECK(gf2t_set_one_slab(ACC.w() + n * pitch, pitch, nw, (int)add, nullptr));
        return GF2T_OK;
    }

    // interval ends: gcds, resolution, retirement and compaction
    int check_intervals() {
        std::vector<size_t> ends;
        for (size_t i = 0; i < live.size(); i++)
            if (live[i].kdone >= live[i].kb) ends.push_back(i);
        if (ends.empty()) return GF2T_OK;
        std::vector<std::vector<u64>> acc(ends.size(), std::vector<u64>(nw));
        for (size_t e = 0; e < ends.size(); e++)
            SV_CHECK(gf2t_memcpy_d2h(acc[e].data(), ACC.w() + ends[e] * pitch, nw * 8, nullptr));
        std::vector<gf2t_scan_result> rows(ends.size());
        std::vector<std::string> hexes(ends.size());
        std::vector<char> finished(ends.size(), 0);
        const int threads = P.gcd_threads > 0 ? P.gcd_threads : 8;
        parallel_for(ends.size(), threads, [&](size_t e) {
            const Slot &q = live[ends[e]];
            Poly g = gcd_fast(trinomial(r, q.s), from_words(acc[e].data(), nw));
            memset(&rows[e], 0, sizeof rows[e]);
            if (deg(g) > 0 && resolve_hit(g, q.ka, q.kb, q.rhigh, rows[e], hexes[e]))
                finished[e] = 1;
            else if (q.kb >= q.rhigh) {
                rows[e].kind = q.limited ? GF2T_SCAN_GAVE_UP : GF2T_SCAN_PRIMITIVE;
                finished[e] = 1;
            }
        });
        std::vector<char> gone(live.size(), 0);
        for (size_t e = 0; e < ends.size(); e++) {
            Slot &q = live[ends[e]];
            if (finished[e]) {
                finish(q, rows[e], hexes[e]);
                gone[ends[e]] = 1;
            } else {
                q.L = std::min(2 * q.L, P.Lmax ? P.Lmax : (u64)4096);
                next_interval(q);
                SV_CHECK(gf2t_set_one_slab(ACC.w() + ends[e] * pitch, pitch, nw, 1, nullptr));
            }
        }
        // compaction: live slots keep their order
        size_t nl = 0;
        for (size_t i = 0; i < live.size(); i++) {
            if (gone[i]) continue;
            if (nl != i) {
                SV_CHECK(gf2t_memcpy_d2d(cur + nl * tab, cur + i * tab, tab * 8, nullptr));
                SV_CHECK(gf2t_memcpy_d2d(ACC.w() + nl * pitch, ACC.w() + i * pitch, nw * 8,
                                         nullptr));
                live[nl] = live[i];
                svals[nl] = svals[i];
            }
            nl++;
        }
        if (nl != live.size()) {
            live.resize(nl);
            svals.resize(nl);
            if (nl) SV_CHECK(gf2t_memcpy_h2d(SV.w(), svals.data(), nl * 8, nullptr));
            sq2_stale = true;
        }
        return GF2T_OK;
    }

    // one round: every live slot folds its next m degrees into ACC
    int round() {
        const int n = (int)live.size();
        if (!n) return GF2T_OK;
        SV_CHECK(gf2t_horner_slab(cur, pitch, m, r, SV.w(), C.w(), pitch, n, nullptr));
        SV_CHECK(gf2t_mulmod_slab(ACC.w(), pitch, C.w(), pitch, 0, r, SV.w(), ACC.w(), pitch, n,
                                  WORK.p, nullptr));
        if (sq2 && sq2_stale) {
            SV_CHECK(gf2t_sq2_slab_plan_set(sq2, svals.data(), (size_t)n, nullptr));
            sq2_stale = false;
        }
        SV_CHECK(gf2t_scan::slab_square_steps(cur, oth, pitch, r, SV.w(), (size_t)n * m, m,
                                              (u64)m, nullptr, sq2));
        for (auto &q : live) q.kdone += (u64)m;
        return GF2T_OK;
    }

    int step(int rounds) {
        for (int i = 0; i < std::max(rounds, 1); i++) {
            SV_CHECK(admit());
            if (live.empty()) break;
            SV_CHECK(round());
            SV_CHECK(check_intervals());
        }
        return GF2T_OK;
    }
};

extern "C" int gf2t_server_create(uint64_t r, const gf2t_scan_params *params, size_t mem_bytes,
                                  gf2t_server **out) {
    if (!out || !params || r < 5 || r >= ((uint64_t)1 << 31)) {
        gf2t::set_error("gf2t_server_create: bad arguments");
        return GF2T_EINVAL;
    }
    *out = nullptr;
    auto *sv = new gf2t_server();
    int rc = sv->create(r, *params, mem_bytes);
    if (rc) {
        delete sv;
        return rc;
    }
    *out = sv;
    return GF2T_OK;
}

extern "C" int gf2t_server_capacity(const gf2t_server *sv, size_t *slots, size_t *bytes_per_slot) {
    if (!sv) return GF2T_EINVAL;
    if (slots) *slots = sv->cap;
    if (bytes_per_slot) *bytes_per_slot = sv->per;
    return GF2T_OK;
}

extern "C" int gf2t_server_submit(gf2t_server *sv, uint64_t job, const uint64_t *svals, size_t n,
                                  uint64_t k0, uint64_t maxd) {
    if (!sv || (n && !svals) || k0 < 1) {
        gf2t::set_error("gf2t_server_submit: bad arguments");
        return GF2T_EINVAL;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    return sv->submit(job, svals, n, k0, maxd);
}

extern "C" int gf2t_server_step(gf2t_server *sv, int rounds, size_t *live, size_t *queued) {
    if (!sv) return GF2T_EINVAL;
    std::lock_guard<std::mutex> lk(sv->mu);
    int rc;
    try {
        rc = sv->step(rounds);
    } catch (const std::exception &e) {
        gf2t::set_error("gf2t_server_step: %s", e.what());
        rc = GF2T_EHIP;
    }
    if (live) *live = sv->live.size();
    if (queued) *queued = sv->queue.size();
    return rc;
}

extern "C" int gf2t_server_poll(gf2t_server *sv, size_t max_rows, uint64_t *jobs, char *buf,
                                size_t cap, size_t *nrows) {
    if (!sv || !nrows || (max_rows && (!jobs || !buf))) return GF2T_EINVAL;
    std::lock_guard<std::mutex> lk(sv->mu);
    size_t used = 0, k = 0;
    if (cap) buf[0] = 0;
    while (k < max_rows && !sv->done.empty()) {
        const Done &d = sv->done.front();
        if (used + d.line.size() + 2 > cap) break;
        memcpy(buf + used, d.line.c_str(), d.line.size());
        used += d.line.size();
        buf[used++] = '\n';
        buf[used] = 0;
        jobs[k++] = d.job;
        sv->done.pop_front();
    }
    *nrows = k;
    return GF2T_OK;
}

extern "C" void gf2t_server_destroy(gf2t_server *sv) { delete sv; }
