// coarse_sieve -- command-line drop-in for the reference's coarse_sieve
// (src/coarse_sieve.cu:1273-1570) on one MI355X.
//   coarse_sieve <r> <depth> <out_prefix> [--smax S] [--found] [--no-cert]
//                [--no-survivors] [--load f.found.bin]
//                [--load-survivors f.txt] [--min-depth D]
//   coarse_sieve <r> <depth> <smax> <out_prefix> ...   (legacy form)
//   coarse_sieve --selftest
// Outputs are byte-compatible: <prefix>.found.cert ("s d p<hex>",
// ascending s), <prefix>.survivors.txt, <prefix>.found.bin (u64 keys for
// s = 1..smax).  Degrees up to 34 run with whole tables resident in
// 288 GB, 35 and 36 with the log table in 2 and 4 value slices, all in
// seconds, so the low-memory bucket mode and its checkpoints are not
// needed here.
#include "gf2t.h"
#include "gf2t_engine.hpp"
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using gf2t::u64;

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static u64 irreducible_count(int d) {
    long long acc = 0;
    for (int k = 1; k <= d; k++) {
        if (d % k) continue;
        int n = k, mu = 1;
        for (int p = 2; p * p <= n; p++)
            if (n % p == 0) {
                n /= p;
                if (n % p == 0) { mu = 0; break; }
                mu = -mu;
            }
        if (mu && n > 1) mu = -mu;
        acc += (long long)mu * (1LL << (d / k));
    }
    return (u64)(acc / d);
}

static int selftest() {
    // r = 136279841, every s <= 999999, degrees 2..20: class counts are
    // the irreducible counts, 913,104 s are factored (src/README.md:86-90)
    // and s = 33 carries "33 5 p29" (README.md:89-93)
    const u64 r = 136279841, smax = 999999;
    std::vector<u64> best;
    std::vector<gf2t_sieve_stats> st;
    double t0 = now_s();
    gf2t::run_gpu(r, 20, smax, best, 2, &st);
    int bad = 0;
    for (int d = 2; d <= 20; d++)
        if (st[d].classes != irreducible_count(d)) {
            fprintf(stderr, "d=%d: %" PRIu64 " classes, expected %" PRIu64 "\n", d,
                    st[d].classes, irreducible_count(d));
            bad++;
        }
    u64 found = 0;
    for (u64 s = 1; s <= smax; s++) found += best[s] != ~0ULL;
    if (found != 913104) { fprintf(stderr, "%" PRIu64 " factored, expected 913104\n", found); bad++; }
    if (best[33] != ((5ULL << 48) | 0x29)) { fprintf(stderr, "s=33 mismatch\n"); bad++; }
    printf("selftest: %s (%.2fs)\n", bad ? "FAIL" : "PASS", now_s() - t0);
    return bad ? 1 : 0;
}

int main(int argc, char **argv) {
    if (argc >= 2 && !strcmp(argv[1], "--selftest")) return selftest();
    if (argc < 4) {
        fprintf(stderr,
                "usage: %s <r> <depth> <out_prefix> [--smax S] [--found] [--no-cert]\n"
                "          [--no-survivors] [--load f.found.bin] [--load-survivors f.txt]\n"
                "          [--min-depth D]\n"
                "       %s --selftest\n", argv[0], argv[0]);
        return 1;
    }
    u64 r = strtoull(argv[1], 0, 10);
    int depth = atoi(argv[2]);
    bool legacy = argc >= 5 && argv[3][0] && strspn(argv[3], "0123456789") == strlen(argv[3]);
    u64 smax = legacy ? strtoull(argv[3], 0, 10) : r / 2;
    const char *prefix = legacy ? argv[4] : argv[3];
    int first = legacy ? 5 : 4;
    bool write_found = false, write_surv = true, write_cert = true;
    const char *load = nullptr, *load_surv = nullptr;
    int min_depth = 2;
    for (int a = first; a < argc; a++) {
        if (!strcmp(argv[a], "--found")) write_found = true;
        else if (!strcmp(argv[a], "--smax") && a + 1 < argc) smax = strtoull(argv[++a], 0, 10);
        else if (!strcmp(argv[a], "--cert")) write_cert = true;
        else if (!strcmp(argv[a], "--no-cert")) write_cert = false;
        else if (!strcmp(argv[a], "--no-survivors")) write_surv = false;
        else if (!strcmp(argv[a], "--load") && a + 1 < argc) load = argv[++a];
        else if (!strcmp(argv[a], "--load-survivors") && a + 1 < argc) load_surv = argv[++a];
        else if (!strcmp(argv[a], "--min-depth") && a + 1 < argc) min_depth = atoi(argv[++a]);
        else fprintf(stderr, "warning: ignoring unknown argument '%s'\n", argv[a]);
    }
    if (min_depth < 2) min_depth = 2;
    if (depth > 36) {
        fprintf(stderr, "depth %d: degrees above 36 are not supported\n", depth);
        return 1;
    }

    gf2t::check(gf2t_init(0), "gf2t_init");
    std::vector<u64> best(smax + 1, ~0ULL);
    if (load) {
        FILE *f = fopen(load, "rb");
        if (!f || fread(best.data() + 1, 8, smax, f) != smax) {
            fprintf(stderr, "%s: cannot read %" PRIu64 " keys\n", load, smax);
            return 1;
        }
        fclose(f);
    } else if (load_surv) {
        // every s not listed has a factor below min_depth: a sentinel key
        // below (min_depth << 48) blocks new writes (src/coarse_sieve.cu:
        // 1449-1477)
        std::fill(best.begin() + 1, best.end(), (u64)(min_depth - 1) << 48);
        FILE *f = fopen(load_surv, "r");
        if (!f) { perror(load_surv); return 1; }
        u64 s;
        while (fscanf(f, "%" SCNu64, &s) == 1)
            if (s >= 1 && s <= smax) best[s] = ~0ULL;
        fclose(f);
    }
    double t0 = now_s();
    std::vector<gf2t_sieve_stats> st;
    gf2t::run_gpu(r, depth, smax, best, min_depth, &st);
    for (int d = min_depth; d <= depth; d++)
        fprintf(stderr, "d=%2d done: %" PRIu64 " classes, %" PRIu64 " marking, %.3fs\n", d,
                st[d].classes, st[d].marking,
                (st[d].ms_table + st[d].ms_classes + st[d].ms_heavy) / 1e3);
    fprintf(stderr, "total sieve time: %.2fs\n", now_s() - t0);
    char path[1024];
    u64 nsurv = 0, ncert = 0;
    for (u64 s = 1; s <= smax; s++) nsurv += best[s] == ~0ULL;
    if (write_surv) {
        snprintf(path, sizeof path, "%s.survivors.txt", prefix);
        FILE *f = fopen(path, "w");
        if (!f) { perror(path); return 1; }
        for (u64 s = 1; s <= smax; s++)
            if (best[s] == ~0ULL) fprintf(f, "%" PRIu64 "\n", s);
        fclose(f);
    }
    fprintf(stderr, "%" PRIu64 " survivors (no factor of degree <= %d)%s%s\n", nsurv, depth,
            write_surv ? " -> " : "", write_surv ? path : "");
    if (write_cert) {
        snprintf(path, sizeof path, "%s.found.cert", prefix);
        FILE *f = fopen(path, "w");
        if (!f) { perror(path); return 1; }
        for (u64 s = 1; s <= smax; s++) {
            u64 b = best[s];
            if (b == ~0ULL || (b & 0xFFFFFFFFFFFFULL) == 0) continue;   // survivor / carried sentinel
            fprintf(f, "%" PRIu64 " %u p%" PRIx64 "\n", s, (unsigned)(b >> 48),
                    (u64)(b & 0xFFFFFFFFFFFFULL));
            ncert++;
        }
        fclose(f);
        fprintf(stderr, "%" PRIu64 " factors -> %s\n", ncert, path);
    }
    if (write_found) {
        snprintf(path, sizeof path, "%s.found.bin", prefix);
        FILE *f = fopen(path, "wb");
        if (!f) { perror(path); return 1; }
        fwrite(best.data() + 1, 8, smax, f);
        fclose(f);
        fprintf(stderr, "packed keys for s=1..%" PRIu64 " -> %s\n", smax, path);
    }
    return 0;
}
