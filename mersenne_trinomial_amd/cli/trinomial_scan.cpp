// trinomial_scan -- command-line drop-in for the reference's deep scan
// (src/trinomial_scan.cu:541-690) on one MI355X: least factor degree in
// (k0, maxd] and the lex-least mask at that degree for every survivor.
//   trinomial_scan <r> <survivors.txt> <k0> <maxd> <out_prefix>
//       [--interval Lmax] [--l0 L0] [--gcd-threads N] [--batch B]
//       [--verbose 0..3]
//   trinomial_scan --selftest
// <out_prefix>.results.txt is appended and resumable (s values already
// listed are skipped):
//     s d p<hex>    least-degree factor, lex-least mask at that degree
//     s rA-B        a factor with degree in [A, B], unresolved
//     s u           no factor of degree <= maxd
//     s primitive   maxd = 0 and no factor up to the Swan bound
// Every survivor of a cohort advances together on the device.  Instead of
// the reference's per-s --state files, <out_prefix>.gf2t.ckpt records the
// degree through which the s values of the running cohort are clear (every
// --ckpt-mins minutes, and on SIGINT / SIGTERM, after which the run writes
// the results it has and exits with status 3); rerunning the same command
// resumes them from there.
#include "gf2t.h"
#include "gf2t_engine.hpp"
#include "scan_ckpt.hpp"
#include <chrono>
#include <map>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

using gf2t::u64;

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// lines of the finished s values (an s the progress callback stopped
// gets none), in the order of sv
static int scan_lines(u64 r, const std::vector<u64> &sv, const gf2t_scan_params &P,
                      std::vector<std::string> &lines) {
    gf2t_scan_results *res = nullptr;
    int rc = gf2t_scan_run(r, sv.data(), sv.size(), &P, &res);
    if (rc != GF2T_OK) return rc;
    lines.clear();
    for (size_t i = 0; i < gf2t_scan_results_count(res); i++) {
        gf2t_scan_result row;
        gf2t_scan_results_get(res, i, &row, nullptr, 0);
        if (row.kind == GF2T_SCAN_STOPPED) continue;
        std::vector<char> buf(row.hex_len + 96);
        rc = gf2t_scan_line(res, i, buf.data(), buf.size());
        if (rc != GF2T_OK) break;
        lines.push_back(buf.data());
    }
    gf2t_scan_results_free(res);
    return rc;
}

static int selftest() {
    // r = 4423: every s <= 2211 scanned from degree 2 to 200; each "p"
    // line's factor must divide T (host gcd of T and the factor).
    const u64 r = 4423;
    std::vector<u64> sv;
    for (u64 s = 1; 2 * s < r; s++) sv.push_back(s);
    gf2t_scan_params P = {};
    P.k0 = 1;
    P.maxd = 200;
    std::vector<std::string> lines;
    double t0 = now_s();
    gf2t::check(scan_lines(r, sv, P, lines), "gf2t_scan_run");
    int bad = 0, nfound = 0, nu = 0;
    for (const std::string &ln : lines) {
        u64 s, d;
        char tag[8192];
        if (sscanf(ln.c_str(), "%" SCNu64 " %" SCNu64 " p%8190s", &s, &d, tag) == 3) {
            nfound++;
            std::vector<u64> f, t((r + 64) / 64, 0);
            size_t n = strlen(tag);
            f.assign((n * 4 + 63) / 64 + 1, 0);
            for (size_t i = 0; i < n; i++) {
                int c = tag[n - 1 - i];
                u64 v = c <= '9' ? c - '0' : (c | 32) - 'a' + 10;
                f[(i * 4) >> 6] |= v << ((i * 4) & 63);
            }
            t[0] |= 1;
            t[s >> 6] |= 1ULL << (s & 63);
            t[r >> 6] |= 1ULL << (r & 63);
            std::vector<u64> g(f.size());
            int64_t gd = gf2t_gcd_host(t.data(), t.size(), f.data(), f.size(), g.data(), g.size());
            if (gd != (int64_t)d) {
                fprintf(stderr, "s=%" PRIu64 ": factor of degree %" PRIu64 " does not divide T\n", s, d);
                bad++;
            }
        } else if (sscanf(ln.c_str(), "%" SCNu64 " %7s", &s, tag) == 2 && !strcmp(tag, "u")) {
            nu++;
        }
    }
    printf("r=4423: %zu results, %d factors, %d u (%.2fs)\n", lines.size(), nfound, nu, now_s() - t0);
    bool ok = bad == 0 && lines.size() == sv.size();
    printf("\nSCAN SELFTEST: %s\n", ok ? "ALL PASS" : "FAIL");
    return ok ? 0 : 1;
}

int main(int argc, char **argv) {
    if (argc >= 2 && !strcmp(argv[1], "--selftest")) {
        gf2t::check(gf2t_init(0), "gf2t_init");
        return selftest();
    }
    if (argc < 6) {
        fprintf(stderr,
                "usage: %s <r> <survivors.txt> <k0> <maxd> <out_prefix>\n"
                "          [--interval Lmax] [--l0 L0] [--gcd-threads N] [--batch B]\n"
                "          [--verbose 0..3] [--ckpt-mins N] [--no-ckpt]\n"
                "       %s --selftest\n"
                "k0 must be the depth the coarse sieve actually reached; maxd 0 scans\n"
                "to the Swan bound.\n", argv[0], argv[0]);
        return 1;
    }
    u64 r = strtoull(argv[1], 0, 10);
    const char *survf = argv[2];
    gf2t_scan_params P = {};
    P.k0 = strtoull(argv[3], 0, 10);
    P.maxd = strtoull(argv[4], 0, 10);
    P.L0 = 64;
    P.Lmax = 4096;
    P.gcd_threads = 8;
    std::string prefix = argv[5];
    double ckpt_mins = 15;
    bool no_ckpt = false;
    long stop_after = -1;
    for (int i = 6; i < argc; i++) {
        if (!strcmp(argv[i], "--ckpt-mins") && i + 1 < argc) ckpt_mins = atof(argv[++i]);
        else if (!strcmp(argv[i], "--no-ckpt")) no_ckpt = true;
        else if (!strcmp(argv[i], "--stop-after-intervals") && i + 1 < argc)   // test hook
            stop_after = atol(argv[++i]);
        else if (!strcmp(argv[i], "--interval") && i + 1 < argc) P.Lmax = strtoull(argv[++i], 0, 10);
        else if (!strcmp(argv[i], "--l0") && i + 1 < argc) P.L0 = strtoull(argv[++i], 0, 10);
        else if (!strcmp(argv[i], "--gcd-threads") && i + 1 < argc) P.gcd_threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--batch") && i + 1 < argc) P.batch = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--verbose") && i + 1 < argc) P.verbose = atoi(argv[++i]);
        else {
            fprintf(stderr, "unknown option %s\n", argv[i]);
            return 1;
        }
    }
    if (P.Lmax < 1) P.Lmax = 1;
    if (P.L0 < 1) P.L0 = 1;
    if (P.L0 > P.Lmax) P.L0 = P.Lmax;
    if (!(P.k0 >= 1 && (P.maxd == 0 || P.maxd > P.k0))) {
        fprintf(stderr, "need 1 <= k0 < maxd (or maxd = 0)\n");
        return 1;
    }
    std::vector<u64> surv;
    {
        FILE *f = fopen(survf, "r");
        if (!f) { perror(survf); return 1; }
        u64 s;
        while (fscanf(f, "%" SCNu64, &s) == 1) {
            if (s >= 1 && 2 * s < r) surv.push_back(s);
            else fprintf(stderr, "skipping s=%" PRIu64 " (need 1 <= s <= (r-1)/2)\n", s);
        }
        fclose(f);
    }
    std::string rpath = prefix + ".results.txt";
    std::set<u64> done;
    scan_ckpt::read_done(rpath, done);
    std::vector<u64> todo;
    for (u64 s : surv)
        if (!done.count(s)) todo.push_back(s);
    FILE *fres = fopen(rpath.c_str(), "a");
    if (!fres) { perror(rpath.c_str()); return 1; }
    fprintf(stderr,
            "trinomial_scan r=%" PRIu64 " k0=%" PRIu64 " maxd=%" PRIu64
            ": %zu survivors (%zu already done), %d gcd threads\n",
            r, P.k0, P.maxd, surv.size(), surv.size() - todo.size(), P.gcd_threads);
    gf2t::check(gf2t_init(0), "gf2t_init");
    // checkpoint: s values of an interrupted cohort resume from the degree
    // they were clear through
    scan_ckpt::Writer W;
    W.p.r = r;
    W.p.skip = P.k0;
    W.p.m = 1;
    W.p.maxd = P.maxd;
    W.every_s = ckpt_mins * 60.0;
    W.stop_after = stop_after;
    W.verbose = P.verbose;
    W.last = scan_ckpt::Writer::now();
    if (!no_ckpt) W.path = prefix + ".gf2t.ckpt";
    std::map<u64, u64> resume;   // s -> degree it is clear through
    {
        scan_ckpt::Ckpt c;
        if (!W.path.empty() && scan_ckpt::load(W.path, c)) {
            if (c.p == W.p) {
                u64 klo = ~0ULL, khi = 0;
                for (const auto &e : c.ent)
                    if (!done.count(e.first)) {
                        u64 &k = resume[e.first];
                        k = std::max(k, e.second);
                        klo = std::min(klo, e.second);
                        khi = std::max(khi, e.second);
                    }
                if (!resume.empty())
                    fprintf(stderr, "resuming %zu s from %s (clear through degree %" PRIu64
                                    "..%" PRIu64 ")\n", resume.size(), W.path.c_str(), klo, khi);
            } else {
                fprintf(stderr, "checkpoint %s is for other parameters; ignored\n", W.path.c_str());
            }
        }
    }
    W.carry = resume;
    P.progress = &scan_ckpt::Writer::cb;
    P.progress_user = &W;
    scan_ckpt::install_signals();
    // results are flushed a chunk at a time so an interrupted run resumes
    const size_t chunk = P.batch > 0 ? (size_t)P.batch : 4096;
    double t0 = now_s();
    u64 nfound = 0, nu = 0, nrange = 0, nprim = 0;
    for (size_t at = 0; at < todo.size() && !W.stopped; at += chunk) {
        std::vector<u64> sv(todo.begin() + at, todo.begin() + std::min(todo.size(), at + chunk));
        // resumed s values first (deepest first), then the rest from k0
        std::map<u64, std::vector<u64>> by_k0;
        for (u64 s : sv) {
            auto it = resume.find(s);
            by_k0[it == resume.end() ? P.k0 : std::max(P.k0, it->second)].push_back(s);
        }
        std::vector<std::string> lines;
        std::vector<u64> finished;
        for (auto g = by_k0.rbegin(); g != by_k0.rend(); ++g) {
            if (W.stopped) break;
            gf2t_scan_params Q = P;
            Q.k0 = g->first;
            std::vector<std::string> part;
            gf2t::check(scan_lines(r, g->second, Q, part), "gf2t_scan_run");
            for (const std::string &ln : part) finished.push_back(strtoull(ln.c_str(), 0, 10));
            lines.insert(lines.end(), part.begin(), part.end());
        }
        for (const std::string &ln : lines) {
            fprintf(fres, "%s\n", ln.c_str());
            if (ln.find(" primitive") != std::string::npos) nprim++;
            else if (ln.find(" p") != std::string::npos) nfound++;
            else if (ln.find(" r") != std::string::npos) nrange++;
            else nu++;
        }
        fflush(fres);
        fsync(fileno(fres));
        W.flushed(finished);
        fprintf(stderr, "%zu/%zu scanned (%.1fs)\n", std::min(todo.size(), at + chunk), todo.size(),
                now_s() - t0);
    }
    if (W.stopped) {
        fclose(fres);
        fprintf(stderr, "interrupted: results so far in %s; run the same command to resume\n",
                rpath.c_str());
        return 3;
    }
    if (!W.path.empty()) remove(W.path.c_str());
    fclose(fres);
    fprintf(stderr, "done: %" PRIu64 " found, %" PRIu64 " ranges, %" PRIu64 " u, %" PRIu64
            " primitive, %.1fs total\n", nfound, nrange, nu, nprim, now_s() - t0);
    return 0;
}
