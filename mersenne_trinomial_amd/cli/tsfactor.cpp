// tsfactor -- command-line drop-in for the reference's blocked-DDF
// scanner (src/tsfactor.cu:2119-2520) on one MI355X.
//   tsfactor <r> <survivors.txt> --skip D [options]
//   tsfactor --selftest [--gcd-threads N]
//   tsfactor --bench <r> [s] [--bench-gcd]
// Results go to <out>.results.txt (default ts), appended and resumable
// (s values already listed are skipped), one line per survivor in file
// order, and to stdout:
//     s d p<hex>    least-degree factor, numerically least at that degree
//     s u           no factor up to --maxd, or -z gave up first
//     s primitive   no factor up to the Swan bound
//     s rA-B        a factor with degree in [A, B], unresolved
// The survivors advance in device cohorts of the blocked scheme (block of
// m degrees, one product per block), so intervals end on block
// boundaries counted from --skip exactly as the reference's do, and a
// factor found in the last block beyond the bound is reported as the
// reference reports it.  -z gives up where the reference's interval
// schedule (linear, or --sched opt) first exceeds Q blocks; -Z, --q0 and
// the schedule otherwise only shape the intervals, which do not change
// any line.  Checkpoints go to <out>.gf2t.ckpt: the degree through which
// each unfinished s is clear, written every --ckpt-mins minutes and on
// SIGINT / SIGTERM (the run then writes the results it has and exits with
// status 3); rerunning the same command resumes them from there.  The
// reference's per-s TSFC checkpoint (<out>.ckpt and its rotations) of an
// in-flight reference run is read too: its s resumes at the start of the
// interval it was in.
#include "gf2t.h"
#include "gf2t_engine.hpp"
#include "scan_ckpt.hpp"
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <set>
#include <string>
#include <vector>

using gf2t::u64;

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static void usage() {
    fprintf(stderr,
            "Usage: tsfactor <r> <survivors.txt> --skip D [options]\n"
            "       tsfactor --selftest [--gcd-threads N]\n"
            "       tsfactor --bench <r> [s] [--bench-gcd]\n"
            "Options:\n"
            "  --skip D        REQUIRED: inputs have no factor of degree <= D\n"
            "  --maxd D        do not scan degrees > D (emit \"s u\")\n"
            "  --m M           block size 2..64 (default: chosen from r)\n"
            "  --q0 N          blocks in the first interval (default 600/m)\n"
            "  -z Q            emit \"s u\" when q would exceed Q\n"
            "  -Z Q            cap q at Q but keep scanning\n"
            "  --sched opt|linear, --sched-G N, --sched-rho R   interval schedule\n"
            "  --out PREFIX    results -> PREFIX.results.txt (default ts)\n"
            "  --gcd-threads N host threads for the interval gcds\n"
            "  --device N      GPU index\n"
            "  --gpu-gcd-min-bits N   gcd products with both operands >= N bits\n"
            "                  run on the GPU (default 2^16)\n"
            "  --legacy-sq     one squaring per launch\n"
            "  --batch B       s values per device cohort (default: by memory)\n"
            "  --ckpt-mins N   checkpoint cadence (default 15); --no-ckpt: none\n"
            "  -v, -vv         verbose\n"
            "Accepted for compatibility, no effect here: --canzass-max D, --pend-max N,\n"
            "  --gpu-gcd, --allow-no-ntl, --selftest-reps N.\n");
    exit(1);
}

// The reference's interval schedule (QSched, src/tsfactor.cu:834-894):
// linear growth q += q0, or the Euler-optimal widths of --sched opt.
struct Sched {
    long q0 = 15, q = 0, m = 1;
    bool opt = false;
    u64 G = 4000;
    double rho_min = 1.25;
    void grow(u64 k_end) {
        if (!opt) {
            q += q0;
            return;
        }
        u64 w = (u64)q * (u64)m;
        if (!w) {
            // first interval: the k1 of least modelled cost, E = sum over
            // intervals of S(k_prev) * (width + G), S(k) = (16/9) / k
            const u64 k0 = k_end;
            const double C = 16.0 / 9.0;
            double bestE = 1e300;
            u64 bestk1 = 0;
            u64 lo = (u64)(1.02 * sqrt((double)k0 * (double)G)) + 1;
            if (lo <= k0) lo = k0 + (u64)m;
            for (int i = 0; i < 96; i++) {
                u64 k1 = lo + (u64)((double)lo * 0.06 * i);
                double E = 0, kp = (double)k0, kc = (double)k1;
                for (int j = 0; j < 200 && kp < 4e18; j++) {
                    E += (C / kp) * ((kc - kp) + (double)G);
                    double nx = kc * kc / kp - (double)G;
                    double wn = nx - kc, wc = kc - kp;
                    if (wn < wc) wn = wc;
                    if (wn < (rho_min - 1.0) * kc) wn = (rho_min - 1.0) * kc;
                    kp = kc;
                    kc = kc + wn;
                    if (kp > 1e14) break;
                }
                if (E < bestE) {
                    bestE = E;
                    bestk1 = k1;
                }
            }
            w = bestk1 > k0 ? bestk1 - k0 : (u64)m;
        } else {
            // next width from the stationarity condition, never shrinking,
            // at least a (rho_min - 1) fraction of the position
            double kd = (double)k_end, wd = (double)w;
            double wn = wd * kd / (kd - wd) - (double)G;
            if (wn < wd) wn = wd;
            if (wn < (rho_min - 1.0) * kd) wn = (rho_min - 1.0) * kd;
            w = (u64)wn;
        }
        q = (long)((w + (u64)m - 1) / (u64)m);
        if (q < 1) q = 1;
    }
};

struct Opts {
    u64 skip = (u64)-1, maxd = 0;
    int m = 0, batch = 0, verbose = 0, gcd_threads = 0;
    long q0 = 0, zq = 0, Zq = 0;
    Sched sched;
};

// Walks the reference's interval loop (src/tsfactor.cu:1266-1333) for a
// degree bound rhigh: returns the last degree it scans, and sets *gave_up
// when -z stops it first.
static u64 scanned_to(const Opts &O, u64 rhigh, bool *gave_up) {
    Sched Q = O.sched;
    Q.q = 0;
    const u64 m = (u64)O.m;
    u64 k = O.skip + 1;
    Q.grow(k);
    if (O.Zq && Q.q > O.Zq) Q.q = O.Zq;
    *gave_up = false;
    for (;;) {
        if (k > rhigh) return k - 1;
        long q = Q.q;
        u64 k2 = k + (u64)q * m - 1;
        if (k2 > rhigh) {
            q = (long)((rhigh - k + m) / m);
            k2 = k + (u64)q * m - 1;
            Q.q = q;
        }
        if (O.zq && q > O.zq) {
            *gave_up = true;
            return k - 1;
        }
        k = k2 + 1;
        Q.grow(k - 1);
        if (O.Zq && Q.q > O.Zq) Q.q = O.Zq;
    }
}

static int scan_lines(u64 r, const std::vector<u64> &sv, const gf2t_scan_params &P,
                      std::vector<std::string> &lines) {
    gf2t_scan_results *res = nullptr;
    int rc = gf2t_scan_run(r, sv.data(), sv.size(), &P, &res);
    if (rc != GF2T_OK) return rc;
    lines.clear();
    for (size_t i = 0; i < gf2t_scan_results_count(res); i++) {
        gf2t_scan_result row;
        gf2t_scan_results_get(res, i, &row, nullptr, 0);
        if (row.kind == GF2T_SCAN_STOPPED) {   // the checkpoint holds it
            lines.push_back(std::string());
            continue;
        }
        std::vector<char> buf(row.hex_len + 96);
        rc = gf2t_scan_line(res, i, buf.data(), buf.size());
        if (rc != GF2T_OK) break;
        lines.push_back(buf.data());
    }
    gf2t_scan_results_free(res);
    return rc;
}

// One line per s of sv, in order: s values are grouped by the bound the
// reference would scan them to (Swan bound or --maxd, cut short by -z).
// s values in `resume` start from the degree recorded there; an empty
// line is an s the progress callback stopped.
static int tsfactor_lines(u64 r, const std::vector<u64> &sv, const Opts &O,
                          const std::map<u64, u64> &resume, scan_ckpt::Writer *W,
                          std::vector<std::string> &out) {
    struct Group {
        u64 k0, maxd;
        bool past;
        std::vector<size_t> idx;
    };
    std::map<std::tuple<u64, u64, int>, Group> groups;
    out.assign(sv.size(), std::string());
    for (size_t i = 0; i < sv.size(); i++) {
        int swan = 0;
        const u64 swan_hi = gf2t_swan_bound(r, sv[i], &swan);
        const bool limited = O.maxd && O.maxd < swan_hi;
        const u64 rhigh = limited ? O.maxd : swan_hi;
        bool cut = false;
        const u64 last = scanned_to(O, rhigh, &cut);
        if (cut && last <= O.skip) {   // the first interval already exceeds -z
            out[i] = std::to_string(sv[i]) + " u";
            continue;
        }
        // cut: degrees up to last (a block boundary), then "u";
        // otherwise the bound with the last block's overshoot
        const u64 maxd = cut ? last : (limited ? O.maxd : 0);
        auto rs = resume.find(sv[i]);
        const u64 k0 = rs == resume.end() ? O.skip : std::max(O.skip, rs->second);
        Group &G = groups[std::make_tuple(k0, maxd, cut ? 0 : 1)];
        G.k0 = k0;
        G.maxd = maxd;
        G.past = !cut;
        G.idx.push_back(i);
    }
    // resumed s (larger k0) first, so an interruption during the fresh
    // cohort does not leave them waiting behind it
    for (auto it = groups.rbegin(); it != groups.rend(); ++it) {
        Group &G = it->second;
        if (W && W->stopped) break;   // the rest restart from the checkpoint or --skip
        gf2t_scan_params P = {};
        P.k0 = G.k0;
        if (W) {
            P.progress = &scan_ckpt::Writer::cb;
            P.progress_user = W;
        }
        P.maxd = G.maxd;
        P.L0 = 64;
        P.Lmax = 4096;
        P.gcd_threads = O.gcd_threads > 0 ? O.gcd_threads : 8;
        P.batch = O.batch;
        P.verbose = O.verbose;
        P.block_m = O.m;
        P.past_bound = G.past ? 1 : 0;
        std::vector<u64> s;
        for (size_t i : G.idx) s.push_back(sv[i]);
        std::vector<std::string> lines;
        int rc = scan_lines(r, s, P, lines);
        if (rc != GF2T_OK) return rc;
        if (lines.size() != s.size()) {
            fprintf(stderr, "tsfactor: %zu results for %zu s values\n", lines.size(), s.size());
            return GF2T_EINTERNAL;
        }
        for (size_t j = 0; j < G.idx.size(); j++) out[G.idx[j]] = lines[j];
    }
    return GF2T_OK;
}

// block size the library picks for r when --m is not given (kept in step
// with the deep scan's own choice)
static int default_m(u64 r) {
    int m = r < (1u << 14) ? 4 : r < (1u << 19) ? 16 : 32;
    while (m > 1 && 2 * (u64)m >= r) m /= 2;
    return m;
}

// SplitMix64 (Steele, Lea and Flood): a generator that is not linear over
// GF(2), so random residues do not collapse remainder sequences
static u64 splitmix(u64 &st) {
    u64 z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static int bench(u64 r, u64 s, bool with_gcd) {
    if (!s) s = (r - 1) / 2;
    const size_t nw = (size_t)((r + 63) >> 6), pitch = (nw + 31) & ~(size_t)31;
    const int m = 24;
    uint64_t *H = nullptr, *H2 = nullptr, *A = nullptr, *Cb = nullptr, *S = nullptr;
    void *W = nullptr;
    gf2t::check(gf2t_malloc((void **)&H, (size_t)m * pitch * 8), "gf2t_malloc");
    gf2t::check(gf2t_malloc((void **)&H2, (size_t)m * pitch * 8), "gf2t_malloc");
    gf2t::check(gf2t_malloc((void **)&A, pitch * 8), "gf2t_malloc");
    gf2t::check(gf2t_malloc((void **)&Cb, pitch * 8), "gf2t_malloc");
    gf2t::check(gf2t_malloc((void **)&S, 8), "gf2t_malloc");
    gf2t::check(gf2t_malloc(&W, gf2t_mulmod_slab_workspace_bytes(r, 1)), "gf2t_malloc");
    gf2t::check(gf2t_memcpy_h2d(S, &s, 8, nullptr), "gf2t_memcpy_h2d");
    gf2t::check(gf2t_ddf_init_slab(H, H2, pitch, m, r, S, 1, 1, nullptr), "gf2t_ddf_init_slab");
    gf2t::check(gf2t_set_one_slab(A, pitch, nw, 1, nullptr), "gf2t_set_one_slab");
    gf2t::check(gf2t_device_sync(), "gf2t_device_sync");
    // one squaring of a whole h-table, a Horner product, one modmul
    double t0 = now_s();
    for (int i = 0; i < 6; i++) {
        gf2t::check(gf2t_square_fold_slab_g(H, H2, pitch, r, S, m, m, nullptr), "square");
        std::swap(H, H2);
    }
    gf2t::check(gf2t_device_sync(), "gf2t_device_sync");
    const double t_sq = (now_s() - t0) / 6 / m;
    t0 = now_s();
    for (int i = 0; i < 4; i++)
        gf2t::check(gf2t_horner_slab(H, pitch, m, r, S, Cb, pitch, 1, nullptr), "horner");
    gf2t::check(gf2t_device_sync(), "gf2t_device_sync");
    const double t_h = (now_s() - t0) / 4;
    gf2t::check(gf2t_mulmod_slab(A, pitch, Cb, pitch, 0, r, S, A, pitch, 1, W, nullptr), "mul");
    gf2t::check(gf2t_device_sync(), "gf2t_device_sync");
    t0 = now_s();
    for (int i = 0; i < 4; i++)
        gf2t::check(gf2t_mulmod_slab(A, pitch, Cb, pitch, 0, r, S, A, pitch, 1, W, nullptr),
                    "mul");
    gf2t::check(gf2t_device_sync(), "gf2t_device_sync");
    const double t_mul = (now_s() - t0) / 4;
    int mo = (int)(sqrt(t_mul / t_sq) + 0.5) & ~1;
    if (mo < 8) mo = 8;
    if (mo > 64) mo = 64;
    printf("r=%" PRIu64 " s=%" PRIu64 ": squaring %.4f ms, Horner (m=%d) %.3f ms, "
           "modmul %.3f ms -> m = %d; per degree at that m %.4f ms\n",
           r, s, t_sq * 1e3, m, t_h * 1e3, t_mul * 1e3, mo,
           (t_sq * mo + (t_h + t_mul) / mo) * 1e3);
    gf2t_free(H);
    gf2t_free(H2);
    gf2t_free(A);
    gf2t_free(Cb);
    gf2t_free(S);
    gf2t_free(W);
    if (with_gcd) {
        // gcd of a random residue against x^r + x^s + 1
        std::vector<u64> T((r + 64) / 64, 0), X(nw, 0);
        T[0] |= 1;
        T[s >> 6] |= 1ULL << (s & 63);
        T[r >> 6] |= 1ULL << (r & 63);
        u64 st = 0x1234abcdULL;
        for (auto &w : X) w = splitmix(st);
        if (r & 63) X[nw - 1] &= (1ULL << (r & 63)) - 1;
        std::vector<u64> g(T.size());
        t0 = now_s();
        int64_t gd = gf2t_gcd_host(X.data(), X.size(), T.data(), T.size(), g.data(), g.size());
        printf("gcd bench (random residue vs T): %.2f s (gdeg %" PRId64 ")\n", now_s() - t0, gd);
    }
    return 0;
}

static int selftest(int gcd_threads) {
    // r = 4423, every s <= 2211 from degree 3 to 300: the blocked scan at
    // m = 8 without the overshoot must give the plain per-degree scan's
    // lines, and every factor must divide its trinomial
    const u64 r = 4423;
    std::vector<u64> sv;
    for (u64 s = 1; 2 * s < r; s++) sv.push_back(s);
    double t0 = now_s();
    gf2t_scan_params P = {};
    P.k0 = 2;
    P.maxd = 300;
    P.gcd_threads = gcd_threads > 0 ? gcd_threads : 8;
    P.block_m = 1;
    std::vector<std::string> plain, blocked;
    gf2t::check(scan_lines(r, sv, P, plain), "gf2t_scan_run");
    P.block_m = 8;
    gf2t::check(scan_lines(r, sv, P, blocked), "gf2t_scan_run");
    int bad = 0, nfound = 0;
    for (size_t i = 0; i < sv.size(); i++) {
        if (plain[i] != blocked[i]) {
            if (bad < 5)
                fprintf(stderr, "s=%" PRIu64 ": plain '%s' blocked '%s'\n", sv[i],
                        plain[i].c_str(), blocked[i].c_str());
            bad++;
            continue;
        }
        u64 s, d;
        char hex[4096];
        if (sscanf(plain[i].c_str(), "%" SCNu64 " %" SCNu64 " p%4094s", &s, &d, hex) != 3) continue;
        nfound++;
        const size_t n = strlen(hex);
        std::vector<u64> f((n * 4 + 63) / 64 + 1, 0), t((r + 64) / 64, 0);
        for (size_t j = 0; j < n; j++) {
            int c = hex[n - 1 - j];
            u64 v = c <= '9' ? (u64)(c - '0') : (u64)((c | 32) - 'a' + 10);
            f[(j * 4) >> 6] |= v << ((j * 4) & 63);
        }
        t[0] |= 1;
        t[s >> 6] |= 1ULL << (s & 63);
        t[r >> 6] |= 1ULL << (r & 63);
        std::vector<u64> g(f.size());
        if (gf2t_gcd_host(t.data(), t.size(), f.data(), f.size(), g.data(), g.size()) !=
            (int64_t)d) {
            fprintf(stderr, "s=%" PRIu64 ": factor of degree %" PRIu64 " does not divide T\n",
                    s, d);
            bad++;
        }
    }
    printf("r=4423: %zu s, %d factors, plain and blocked (m=8) scans %s (%.2fs)\n", sv.size(),
           nfound, bad ? "DIFFER" : "agree", now_s() - t0);
    printf("\nTSFACTOR SELFTEST: %s\n", bad ? "FAIL" : "ALL PASS");
    return bad ? 1 : 0;
}

int main(int argc, char **argv) {
    Opts O;
    bool do_selftest = false, do_bench = false, bench_gcd = false, legacy = false;
    int device = 0;
    u64 gcd_bits = 0;
    double ckpt_mins = 15;
    bool no_ckpt = false;
    long stop_after = -1;
    std::string out_prefix = "ts";
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto need = [&](const char *o) -> const char * {
            if (i + 1 >= argc) {
                fprintf(stderr, "%s needs a value\n", o);
                exit(1);
            }
            return argv[++i];
        };
        if (a == "--selftest") do_selftest = true;
        else if (a == "--bench") do_bench = true;
        else if (a == "--bench-gcd") bench_gcd = true;
        else if (a == "--skip") O.skip = strtoull(need("--skip"), 0, 10);
        else if (a == "--maxd") O.maxd = strtoull(need("--maxd"), 0, 10);
        else if (a == "--m") O.m = atoi(need("--m"));
        else if (a == "--q0") O.q0 = atol(need("--q0"));
        else if (a == "-z") O.zq = atol(need("-z"));
        else if (a == "-Z") O.Zq = atol(need("-Z"));
        else if (a == "--out") out_prefix = need("--out");
        else if (a == "--gcd-threads") O.gcd_threads = atoi(need("--gcd-threads"));
        else if (a == "--device") device = atoi(need("--device"));
        else if (a == "--gpu-gcd-min-bits") gcd_bits = strtoull(need("--gpu-gcd-min-bits"), 0, 10);
        else if (a == "--legacy-sq") legacy = true;
        else if (a == "--batch") O.batch = atoi(need("--batch"));
        else if (a == "--sched") {
            std::string v = need("--sched");
            if (v == "opt") O.sched.opt = true;
            else if (v == "linear") O.sched.opt = false;
            else {
                fprintf(stderr, "bad --sched %s (opt|linear)\n", v.c_str());
                exit(1);
            }
        } else if (a == "--sched-G") {
            O.sched.G = strtoull(need("--sched-G"), 0, 10);
            if (!O.sched.G) O.sched.G = 4000;
        } else if (a == "--sched-rho") {
            double rho = atof(need("--sched-rho"));
            if (rho > 1.0) O.sched.rho_min = rho;
        } else if (a == "--ckpt-mins") ckpt_mins = atof(need("--ckpt-mins"));
        else if (a == "--no-ckpt") no_ckpt = true;
        else if (a == "--stop-after-intervals") stop_after = atol(need(a.c_str()));   // test hook
        else if (a == "--canzass-max" || a == "--pend-max" || a == "--selftest-reps")
            need(a.c_str());
        else if (a == "--gpu-gcd" || a == "--allow-no-ntl") {
        } else if (a == "-v") O.verbose++;
        else if (a == "-vv") O.verbose += 2;
        else if (a.size() > 1 && a[0] == '-' && !isdigit((unsigned char)a[1])) {
            fprintf(stderr, "unknown option %s\n", a.c_str());
            usage();
        } else pos.push_back(a);
    }
    if (do_selftest) {
        gf2t::check(gf2t_init(device), "gf2t_init");
        return selftest(O.gcd_threads);
    }
    if (do_bench) {
        if (pos.empty()) usage();
        const u64 r = strtoull(pos[0].c_str(), 0, 10);
        const u64 s = pos.size() > 1 ? strtoull(pos[1].c_str(), 0, 10) : 0;
        if (r < 5 || (s && 2 * s >= r)) usage();
        gf2t::check(gf2t_init(device), "gf2t_init");
        return bench(r, s, bench_gcd);
    }
    if (pos.size() != 2) usage();
    const u64 r = strtoull(pos[0].c_str(), 0, 10);
    if (!(r & 1) || r < 5) {
        fprintf(stderr, "r must be odd, >= 5\n");
        return 1;
    }
    if (O.skip == (u64)-1) {
        fprintf(stderr, "--skip is REQUIRED (the depth the coarse sieve reached);\n"
                        "scanning starts at skip+1 and factors of degree <= skip would be "
                        "MISSED.\n");
        return 1;
    }
    if (O.skip < 1) {
        fprintf(stderr, "--skip must be >= 1\n");
        return 1;
    }
    if (O.zq && O.Zq) {
        fprintf(stderr, "-z and -Z are mutually exclusive\n");
        return 1;
    }
    const bool m_given = O.m != 0;
    const std::string res_path = out_prefix + ".results.txt";
    std::set<u64> done;
    if (FILE *f = fopen(res_path.c_str(), "r")) {
        fclose(f);
        scan_ckpt::read_done(res_path, done);
        fprintf(stderr, "%zu already completed in %s\n", done.size(), res_path.c_str());
    }
    // the reference's own checkpoint of an in-flight run: its s resumes at
    // the start of the interval it was in (every degree below is clear)
    std::map<u64, u64> resume;
    if (!no_ckpt) {
        scan_ckpt::Tsfc T;
        std::string which;
        if (scan_ckpt::load_tsfc(out_prefix + ".ckpt", T, &which)) {
            if (T.r != r || T.skip != O.skip || done.count(T.s)) {
                fprintf(stderr, "checkpoint %s ignored (r/skip mismatch or s already done)\n",
                        which.c_str());
            } else if (m_given && (u64)O.m != T.m) {
                fprintf(stderr, "checkpoint %s is for m=%" PRIu64 ", not --m %d; ignored\n",
                        which.c_str(), T.m, O.m);
            } else {
                O.m = (int)T.m;
                resume[T.s] = T.k_iv - 1;
                fprintf(stderr, "resuming s=%" PRIu64 " from %s (clear through degree %" PRIu64
                                ", m=%d)\n", T.s, which.c_str(), T.k_iv - 1, O.m);
            }
        }
    }
    if (!O.m) O.m = default_m(r);
    if (O.m < 2 || O.m > 64 || 2 * (u64)O.m >= r) {
        fprintf(stderr, "--m must be in 2..64 with 2m < r\n");
        return 1;
    }
    if (O.q0 <= 0) O.q0 = (600 + O.m - 1) / O.m;
    O.sched.q0 = O.q0;
    O.sched.m = O.m;

    std::vector<u64> svs;
    {
        FILE *f = fopen(pos[1].c_str(), "r");
        if (!f) {
            fprintf(stderr, "cannot open %s\n", pos[1].c_str());
            return 1;
        }
        char line[256];
        while (fgets(line, sizeof line, f)) {
            char *p = line;
            while (*p == ' ' || *p == '\t') p++;
            if (*p == '#' || *p == '\n' || !*p) continue;
            u64 s = strtoull(p, 0, 10);
            if (!(s >= 1 && 2 * s < r)) {
                fprintf(stderr, "bad survivor s=%" PRIu64 " (need 1 <= s <= (r-1)/2)\n", s);
                fclose(f);
                return 1;
            }
            svs.push_back(s);
        }
        fclose(f);
    }
    fprintf(stderr, "%zu survivors from %s\n", svs.size(), pos[1].c_str());
    std::vector<u64> todo;
    for (u64 s : svs)
        if (!done.count(s)) todo.push_back(s);

    gf2t::check(gf2t_init(device), "gf2t_init");
    if (gcd_bits) gf2t::check(gf2t_tune("gcd_dev_bits", (int64_t)gcd_bits), "gf2t_tune");
    if (legacy) gf2t::check(gf2t_tune("sq_fuse", 0), "gf2t_tune");
    fprintf(stderr, "r=%" PRIu64 " skip=%" PRIu64 " m=%d q0=%ld%s%s%s\n", r, O.skip, O.m, O.q0,
            O.zq ? " -z" : "", O.Zq ? " -Z" : "", O.sched.opt ? " sched=opt" : "");
    scan_ckpt::Writer W;
    W.p.r = r;
    W.p.skip = O.skip;
    W.p.m = (u64)O.m;
    W.p.maxd = O.maxd;
    W.p.zq = (u64)O.zq;
    W.every_s = ckpt_mins * 60.0;
    W.stop_after = stop_after;
    W.verbose = O.verbose;
    W.last = scan_ckpt::Writer::now();
    if (!no_ckpt) W.path = out_prefix + ".gf2t.ckpt";
    {
        scan_ckpt::Ckpt c;
        if (!W.path.empty() && scan_ckpt::load(W.path, c)) {
            if (c.p == W.p) {
                size_t nres = 0;
                u64 klo = ~0ULL, khi = 0;
                for (const auto &e : c.ent)
                    if (!done.count(e.first)) {
                        u64 &k = resume[e.first];
                        k = std::max(k, e.second);
                        nres++;
                        klo = std::min(klo, e.second);
                        khi = std::max(khi, e.second);
                    }
                if (nres)
                    fprintf(stderr, "resuming %zu s from %s (clear through degree %" PRIu64
                                    "..%" PRIu64 ")\n", nres, W.path.c_str(), klo, khi);
            } else {
                fprintf(stderr, "checkpoint %s is for other parameters; ignored\n", W.path.c_str());
            }
        }
    }
    W.carry = resume;
    scan_ckpt::install_signals();
    FILE *rf = fopen(res_path.c_str(), "a");
    if (!rf) {
        fprintf(stderr, "cannot append %s\n", res_path.c_str());
        return 1;
    }
    const size_t chunk = O.batch > 0 ? (size_t)O.batch : 4096;
    const double t0 = now_s();
    u64 nfound = 0, nu = 0, nrange = 0, nprim = 0;
    for (size_t at = 0; at < todo.size() && !W.stopped; at += chunk) {
        std::vector<u64> sv(todo.begin() + at, todo.begin() + std::min(todo.size(), at + chunk));
        std::vector<std::string> lines;
        gf2t::check(tsfactor_lines(r, sv, O, resume, &W, lines), "tsfactor scan");
        std::vector<u64> finished;
        for (size_t i = 0; i < lines.size(); i++) {
            const std::string &ln = lines[i];
            if (ln.empty()) continue;   // stopped before it finished
            finished.push_back(sv[i]);
            fprintf(rf, "%s\n", ln.c_str());
            printf("%s\n", ln.c_str());
            if (ln.find(" primitive") != std::string::npos) {
                nprim++;
                fprintf(stderr, "\n*** PRIMITIVE TRINOMIAL CANDIDATE: s=%" PRIu64
                                " -- verify independently! ***\n\n", sv[i]);
            } else if (ln.find(" p") != std::string::npos) nfound++;
            else if (ln.find(" r") != std::string::npos) nrange++;
            else nu++;
        }
        fflush(rf);
        fsync(fileno(rf));
        W.flushed(finished);
        fflush(stdout);
        if (O.verbose)
            fprintf(stderr, "%zu/%zu scanned (%.1fs)\n", std::min(todo.size(), at + chunk),
                    todo.size(), now_s() - t0);
    }
    fclose(rf);
    if (W.stopped) {
        fprintf(stderr, "interrupted: %" PRIu64 " found, %" PRIu64 " u so far in %s; run the "
                        "same command to resume\n", nfound, nu, res_path.c_str());
        return 3;
    }
    if (!W.path.empty()) remove(W.path.c_str());
    fprintf(stderr, "done: %" PRIu64 " found, %" PRIu64 " ranges, %" PRIu64 " u, %" PRIu64
            " primitive, %.1fs\n", nfound, nrange, nu, nprim, now_s() - t0);
    return 0;
}
