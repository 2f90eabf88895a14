// scan_ckpt.hpp -- checkpoint / resume of the deep-scan command-line
// drivers (trinomial_scan, tsfactor).
//
// The scan reports, after every interval of a cohort, the degree k_done
// through which every s still being scanned is clear; restarting those s
// with k0 = k_done continues the search exactly (no factor of degree
// <= k_done exists).  So a checkpoint is the list of those s plus k_done
// and the run parameters it belongs to -- the reference's per-s TSFC
// checkpoint (src/tsfactor.cu:717-828) reduced to what resuming needs.
// It is written atomically (temporary file, fsync, rename), carries a
// CRC-32, and is written every --ckpt-mins minutes and when SIGINT or
// SIGTERM asks the run to stop; a stopped run exits after writing the
// results of every s it finished.
//
// File: "GF2TSCK1", then u64 fields version, r, skip, m, maxd, zq, k_done,
// n, the n s values, and the CRC-32 of everything before it.
#pragma once
#include "gf2t.h"
#include <chrono>
#include <csignal>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

namespace scan_ckpt {

using u64 = uint64_t;

// CRC-32 (reflected, polynomial 0xEDB88320, initial and final xor ~0)
inline uint32_t crc32(const void *buf, size_t len) {
    static uint32_t tab[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            tab[i] = c;
        }
        init = true;
    }
    const unsigned char *p = (const unsigned char *)buf;
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < len; i++) c = tab[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

struct Params {   // what a checkpoint must match to be resumed
    u64 r = 0, skip = 0, m = 0, maxd = 0, zq = 0;
    bool operator==(const Params &o) const {
        return r == o.r && skip == o.skip && m == o.m && maxd == o.maxd && zq == o.zq;
    }
};

struct Ckpt {
    Params p;
    u64 k_done = 0;
    std::vector<u64> s;
};

inline bool save(const std::string &path, const Ckpt &c) {
    std::vector<unsigned char> b(8);
    memcpy(b.data(), "GF2TSCK1", 8);
    auto put = [&](u64 v) {
        unsigned char w[8];
        for (int i = 0; i < 8; i++) w[i] = (unsigned char)(v >> (8 * i));
        b.insert(b.end(), w, w + 8);
    };
    put(1);
    put(c.p.r);
    put(c.p.skip);
    put(c.p.m);
    put(c.p.maxd);
    put(c.p.zq);
    put(c.k_done);
    put(c.s.size());
    for (u64 v : c.s) put(v);
    const uint32_t crc = crc32(b.data(), b.size());
    for (int i = 0; i < 4; i++) b.push_back((unsigned char)(crc >> (8 * i)));
    const std::string tmp = path + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return false;
    bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
    ok = ok && fflush(f) == 0 && fsync(fileno(f)) == 0;
    fclose(f);
    if (!ok || rename(tmp.c_str(), path.c_str()) != 0) {
        remove(tmp.c_str());
        return false;
    }
    return true;
}

inline bool load(const std::string &path, Ckpt &c) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    std::vector<unsigned char> b;
    unsigned char buf[65536];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + got);
    fclose(f);
    if (b.size() < 8 + 8 * 8 + 4 || memcmp(b.data(), "GF2TSCK1", 8)) return false;
    const size_t body = b.size() - 4;
    uint32_t crc = 0;
    for (int i = 0; i < 4; i++) crc |= (uint32_t)b[body + i] << (8 * i);
    if (crc != crc32(b.data(), body)) return false;
    size_t at = 8;
    auto get = [&]() {
        u64 v = 0;
        for (int i = 0; i < 8; i++) v |= (u64)b[at + i] << (8 * i);
        at += 8;
        return v;
    };
    if (get() != 1) return false;
    c.p.r = get();
    c.p.skip = get();
    c.p.m = get();
    c.p.maxd = get();
    c.p.zq = get();
    c.k_done = get();
    const u64 n = get();
    if (body != at + 8 * n) return false;
    c.s.resize(n);
    for (u64 i = 0; i < n; i++) c.s[i] = get();
    return true;
}

// ---- the run's side of the progress callback
inline volatile sig_atomic_t &stop_flag() {
    static volatile sig_atomic_t f = 0;
    return f;
}
inline void on_signal(int) { stop_flag() = 1; }
inline void install_signals() {
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_handler = on_signal;
    sigaction(SIGINT, &sa, nullptr);
    sigaction(SIGTERM, &sa, nullptr);
}

struct Writer {
    std::string path;        // empty: no checkpoints
    Params p;
    double every_s = 15 * 60.0;
    long stop_after = -1;    // test hook: stop at this progress call
    long calls = 0;
    double last = 0;
    bool stopped = false;
    int verbose = 0;
    static double now() {
        using namespace std::chrono;
        return duration<double>(steady_clock::now().time_since_epoch()).count();
    }
    // gf2t_scan_params.progress
    static int cb(void *user, uint64_t k_done, const uint64_t *svals, size_t n) {
        Writer &W = *(Writer *)user;
        const bool hook = W.stop_after >= 0 && ++W.calls >= W.stop_after;
        const bool stop = stop_flag() || hook;
        if (!W.path.empty() && (stop || now() - W.last >= W.every_s)) {
            Ckpt c;
            c.p = W.p;
            c.k_done = k_done;
            c.s.assign(svals, svals + n);
            if (!save(W.path, c)) fprintf(stderr, "checkpoint %s: write failed\n", W.path.c_str());
            else if (W.verbose || stop)
                fprintf(stderr, "checkpoint %s: %zu s clear through degree %llu\n", W.path.c_str(),
                        n, (unsigned long long)k_done);
            W.last = now();
        }
        if (stop) W.stopped = true;
        return stop ? 1 : 0;
    }
};

}  // namespace scan_ckpt
