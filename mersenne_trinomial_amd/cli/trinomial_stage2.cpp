// trinomial_stage2 -- command-line drop-in for the reference's
// trinomial_stage2 (src/trinomial_stage2.cu:1-311) on one MI355X:
// Frobenius iterates F_k = x^(2^k) mod T and the interval product
// ACC = prod_{k=k0..k1} (F_k + x) mod T, with T = x^r + x^s + 1.
//   trinomial_stage2 selftest
//   trinomial_stage2 run <r> <s> <k0> <k1> <out_prefix> [state_file]
//        <out_prefix>.acc / .fk: ceil(r/64) LE u64 words of ACC / F_k1
//        <out_prefix>.meta: "r s k0 k1"; state_file (same layout as the
//        reference's: magic 0x53544732, r, s, k, A words, ACC words) is
//        rewritten every 1024 squarings and resumed from when present.
//   trinomial_stage2 bench <r> <s>
#include "gf2t.h"
#include "gf2t_engine.hpp"
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using gf2t::u64;

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static std::vector<u64> d2h(const u64 *d, size_t nw) {
    std::vector<u64> h(nw);
    gf2t::check(gf2t_memcpy_d2h(h.data(), d, nw * 8, nullptr), "d2h");
    return h;
}

// Selftest model of the ring: bit vectors, x^e reduced as x^(e-r) + x^(e-r+s)
// one bit at a time from the top.
static void naive_reduce(std::vector<uint8_t> &v, u64 r, u64 s) {
    for (size_t e = v.size(); e-- > r;)
        if (v[e]) {
            v[e] = 0;
            v[e - r] ^= 1;
            v[e - r + s] ^= 1;
        }
    v.resize(r);
}
static std::vector<uint8_t> naive_mulmod(const std::vector<uint8_t> &a,
                                         const std::vector<uint8_t> &b, u64 r, u64 s) {
    std::vector<uint8_t> p(2 * r, 0);
    for (u64 i = 0; i < r; i++)
        if (a[i])
            for (u64 j = 0; j < r; j++) p[i + j] ^= b[j];
    naive_reduce(p, r, s);
    return p;
}
static std::vector<u64> to_words(const std::vector<uint8_t> &v) {
    std::vector<u64> w((v.size() + 63) / 64, 0);
    for (size_t i = 0; i < v.size(); i++)
        if (v[i]) w[i >> 6] |= 1ULL << (i & 63);
    return w;
}

static int selftest() {
    int bad = 0;
    {
        // r = 31, s = 3: x^(2^31) = x since x^31 + x^3 + 1 is primitive
        gf2t::TrinGPU S;
        S.init(31);
        S.set_s(3);
        S.set_x();
        u64 oracle = 2, f31 = 0;
        bool ok = true;
        for (int k = 1; k <= 40; k++) {
            S.square_A();
            u64 sq = 0;   // spread, then fold x^e -> x^(e-31) + x^(e-28)
            for (int i = 0; i < 31; i++)
                if ((oracle >> i) & 1) sq |= 1ULL << (2 * i);
            while (sq >> 31) {
                u64 hi = sq >> 31;
                sq = (sq & 0x7FFFFFFFULL) ^ hi ^ (hi << 3);
            }
            oracle = sq;
            u64 got = d2h(S.dA, 1)[0];
            if (got != oracle) ok = false;
            if (k == 31) f31 = got;
        }
        printf("r=31 s=3: squarings vs u64 model k=1..40            %s\n", ok ? "PASS" : "FAIL");
        printf("r=31 s=3: Frobenius order (F_31 == x)               %s\n", f31 == 2 ? "PASS" : "FAIL");
        bad += !ok + (f31 != 2);
        S.fini();
    }
    {
        const u64 r = 4423, s = 1200;
        gf2t::TrinGPU S;
        S.init(r);
        S.set_s(s);
        S.set_x();
        S.reset_acc();
        std::vector<uint8_t> A(r, 0), ACC(r, 0);
        A[1] = 1;
        ACC[0] = 1;
        bool ok = true;
        for (int k = 1; k <= 24; k++) {
            S.square_A();
            std::vector<uint8_t> sq(2 * r, 0);
            for (u64 i = 0; i < r; i++) sq[2 * i] = A[i];
            naive_reduce(sq, r, s);
            A = sq;
            if (d2h(S.dA, S.nw) != to_words(A)) ok = false;
            if (k >= 5 && k <= 12) {
                S.accumulate();
                std::vector<uint8_t> t = A;
                t[1] ^= 1;
                ACC = naive_mulmod(ACC, t, r, s);
            }
        }
        bool okp = d2h(S.dACC, S.nw) == to_words(ACC);
        printf("r=4423 s=1200: squarings vs bit model k=1..24      %s\n", ok ? "PASS" : "FAIL");
        printf("r=4423 s=1200: interval product k=5..12 vs model   %s\n", okp ? "PASS" : "FAIL");
        bad += !ok + !okp;
        S.fini();
    }
    printf(bad ? "\nSTAGE2 SELFTEST: %d FAILURES\n" : "\nSTAGE2 SELFTEST: ALL PASS\n", bad);
    return bad ? 1 : 0;
}

static int bench(u64 r, u64 s) {
    gf2t::TrinGPU S;
    S.init(r);
    S.set_s(s);
    S.set_x();
    for (int k = 0; k < 8; k++) S.square_A();
    // spread-square and FFT-square of the same element must agree
    S.mul_dev(S.dA, S.dA);
    S.reduce(S.dACC, 2 * r - 1);
    S.square_A();
    bool same = d2h(S.dA, S.nw) == d2h(S.dACC, S.nw);
    printf("cross-check spread-square == FFT-square: %s\n", same ? "PASS" : "FAIL");
    void *e0, *e1;
    gf2t::check(gf2t_event_create(&e0), "event");
    gf2t::check(gf2t_event_create(&e1), "event");
    float ms = 0;
    gf2t::check(gf2t_event_record(e0, nullptr), "record");
    for (int k = 0; k < 10; k++) S.square_A();
    gf2t::check(gf2t_event_record(e1, nullptr), "record");
    gf2t::check(gf2t_event_elapsed_ms(e0, e1, &ms), "elapsed");
    printf("squaring+reduction: %.3f ms each\n", ms / 10);
    gf2t::check(gf2t_event_record(e0, nullptr), "record");
    for (int k = 0; k < 3; k++) S.accumulate();
    gf2t::check(gf2t_event_record(e1, nullptr), "record");
    gf2t::check(gf2t_event_elapsed_ms(e0, e1, &ms), "elapsed");
    printf("product+reduction:  %.3f ms each\n", ms / 3);
    gf2t_event_destroy(e0);
    gf2t_event_destroy(e1);
    S.fini();
    return same ? 0 : 1;
}

static bool write_words(const char *path, const std::vector<u64> &w) {
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); return false; }
    bool ok = fwrite(w.data(), 8, w.size(), f) == w.size();
    return fclose(f) == 0 && ok;
}

static int run(u64 r, u64 s, u64 k0, u64 k1, const char *prefix, const char *state_path) {
    const u64 MAGIC = 0x53544732ULL;
    gf2t::TrinGPU S;
    S.init(r);
    S.set_s(s);
    u64 k_start = 0;
    bool resumed = false;
    if (state_path) {
        if (FILE *sf = fopen(state_path, "rb")) {
            u64 hdr[4];
            std::vector<u64> a(S.nw), c(S.nw);
            if (fread(hdr, 8, 4, sf) == 4 && hdr[0] == MAGIC && hdr[1] == r && hdr[2] == s &&
                fread(a.data(), 8, S.nw, sf) == S.nw && fread(c.data(), 8, S.nw, sf) == S.nw) {
                gf2t::check(gf2t_memcpy_h2d(S.dA, a.data(), S.nw * 8, nullptr), "h2d");
                gf2t::check(gf2t_memcpy_h2d(S.dACC, c.data(), S.nw * 8, nullptr), "h2d");
                k_start = hdr[3];
                resumed = true;
                fprintf(stderr, "resumed at k=%" PRIu64 "\n", k_start);
            }
            fclose(sf);
        }
    }
    if (!resumed) {
        S.set_x();
        S.reset_acc();
    }
    double t0 = now_s();
    for (u64 k = k_start + 1; k <= k1; k++) {
        S.square_A();
        if (k >= k0) S.accumulate();
        if (state_path && (k % 1024 == 0 || k == k1)) {
            std::vector<u64> hdr = {MAGIC, r, s, k}, a = d2h(S.dA, S.nw), c = d2h(S.dACC, S.nw);
            std::string tmp = std::string(state_path) + ".tmp";
            FILE *sf = fopen(tmp.c_str(), "wb");
            if (!sf) { perror(tmp.c_str()); return 1; }
            bool ok = fwrite(hdr.data(), 8, 4, sf) == 4 && fwrite(a.data(), 8, S.nw, sf) == S.nw &&
                      fwrite(c.data(), 8, S.nw, sf) == S.nw;
            if (fclose(sf) != 0 || !ok || rename(tmp.c_str(), state_path) != 0) {
                fprintf(stderr, "cannot write %s\n", state_path);
                return 1;
            }
        }
        if (k % 4096 == 0)
            fprintf(stderr, "k=%" PRIu64 " (%.1f sq/s)\n", k, (k - k_start) / (now_s() - t0));
    }
    gf2t::check(gf2t_device_sync(), "synchronize");
    char path[1024];
    snprintf(path, sizeof path, "%s.acc", prefix);
    if (!write_words(path, d2h(S.dACC, S.nw))) return 1;
    snprintf(path, sizeof path, "%s.fk", prefix);
    if (!write_words(path, d2h(S.dA, S.nw))) return 1;
    snprintf(path, sizeof path, "%s.meta", prefix);
    FILE *f = fopen(path, "w");
    if (!f) { perror(path); return 1; }
    fprintf(f, "%" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64 "\n", r, s, k0, k1);
    fclose(f);
    fprintf(stderr, "done: %" PRIu64 " squarings, %.1fs; ACC and F_k1 written for the gcd\n",
            k1 - k_start, now_s() - t0);
    S.fini();
    return 0;
}

int main(int argc, char **argv) {
    const char *mode = argc > 1 ? argv[1] : "selftest";
    if (!strcmp(mode, "selftest")) {
        gf2t::check(gf2t_init(0), "gf2t_init");
        return selftest();
    }
    if (!strcmp(mode, "bench") && argc >= 4) {
        gf2t::check(gf2t_init(0), "gf2t_init");
        return bench(strtoull(argv[2], 0, 10), strtoull(argv[3], 0, 10));
    }
    if (!strcmp(mode, "run") && argc >= 7) {
        gf2t::check(gf2t_init(0), "gf2t_init");
        return run(strtoull(argv[2], 0, 10), strtoull(argv[3], 0, 10), strtoull(argv[4], 0, 10),
                   strtoull(argv[5], 0, 10), argv[6], argc > 7 ? argv[7] : nullptr);
    }
    fprintf(stderr,
            "usage: %s selftest\n"
            "       %s run <r> <s> <k0> <k1> <out_prefix> [state_file]\n"
            "       %s bench <r> <s>\n", argv[0], argv[0], argv[0]);
    return 1;
}
