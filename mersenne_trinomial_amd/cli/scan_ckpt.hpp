// scan_ckpt.hpp -- checkpoint / resume of the deep-scan command-line
// drivers (trinomial_scan, tsfactor).
//
// The scan reports, after every interval of a cohort, the degree k_done
// through which every s still being scanned is clear; restarting those s
// with k0 = k_done continues the search exactly (no factor of degree
// <= k_done exists).  So a checkpoint is a list of (s, k_done) pairs plus
// the run parameters it belongs to -- the reference's per-s TSFC
// checkpoint (src/tsfactor.cu:717-828) reduced to what resuming needs,
// for every unfinished s at once.  It is written atomically (temporary
// file, fsync, rename), carries a CRC-32, and is written every
// --ckpt-mins minutes and when SIGINT or SIGTERM asks the run to stop; a
// stopped run exits after writing the results of every s it finished.
//
// File: "GF2TSCK1", u64 version (2), r, skip, m, maxd, zq, n, then n
// pairs (s, k_done), then the CRC-32 of everything before it.  Version 1
// files (one k_done for all s: version, r, skip, m, maxd, zq, k_done, n,
// the n s values, CRC) are still read.
//
// The reference's own checkpoint (<out>.ckpt, rotated to .ckpt.1 and
// .ckpt.2) is read by load_tsfc: its header names one s and the start
// k_iv of the interval it was in, every degree below k_iv verified clear,
// so that s resumes here with k_done = k_iv - 1.
#pragma once
#include "gf2t.h"
#include <algorithm>
#include <chrono>
#include <csignal>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <unistd.h>
#include <utility>
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
    std::vector<std::pair<u64, u64>> ent;   // (s, degree it is clear through)
};

inline bool read_file(const std::string &path, std::vector<unsigned char> &b) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    b.clear();
    unsigned char buf[65536];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + got);
    fclose(f);
    return true;
}

inline bool save(const std::string &path, const Ckpt &c) {
    std::vector<unsigned char> b(8);
    memcpy(b.data(), "GF2TSCK1", 8);
    auto put = [&](u64 v) {
        unsigned char w[8];
        for (int i = 0; i < 8; i++) w[i] = (unsigned char)(v >> (8 * i));
        b.insert(b.end(), w, w + 8);
    };
    put(2);
    put(c.p.r);
    put(c.p.skip);
    put(c.p.m);
    put(c.p.maxd);
    put(c.p.zq);
    put(c.ent.size());
    for (const auto &e : c.ent) {
        put(e.first);
        put(e.second);
    }
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
    std::vector<unsigned char> b;
    if (!read_file(path, b)) return false;
    if (b.size() < 8 + 7 * 8 + 4 || memcmp(b.data(), "GF2TSCK1", 8)) return false;
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
    const u64 ver = get();
    if (ver != 1 && ver != 2) return false;
    c.p.r = get();
    c.p.skip = get();
    c.p.m = get();
    c.p.maxd = get();
    c.p.zq = get();
    c.ent.clear();
    if (ver == 1) {
        if (body < at + 16) return false;
        const u64 k_done = get();
        const u64 n = get();
        if (n > (body - at) / 8 || body != at + 8 * n) return false;
        for (u64 i = 0; i < n; i++) c.ent.emplace_back(get(), k_done);
        return true;
    }
    const u64 n = get();
    if (n > (body - at) / 16 || body != at + 16 * n) return false;
    for (u64 i = 0; i < n; i++) {
        const u64 s = get();
        c.ent.emplace_back(s, get());
    }
    return true;
}

// The s values already listed in a results file: every line starts with
// its s ("s d p<hex>", "s u", ...).  Lines are read whole, however long
// (a factor line at degree 44,016 is over 11,000 characters), and s is
// taken only at the start of a line.
inline void read_done(const std::string &path, std::set<u64> &done) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return;
    char *line = nullptr;
    size_t cap = 0;
    while (getline(&line, &cap, f) > 0) {
        const char *p = line;
        if (*p < '0' || *p > '9') continue;
        char *end = nullptr;
        const u64 s = strtoull(p, &end, 10);
        if (s && (*end == ' ' || *end == '\t' || *end == '\n' || *end == 0)) done.insert(s);
    }
    free(line);
    fclose(f);
}

// ---- the reference's tsfactor checkpoint (TSFC)
// Its header is the reference's struct CkptHdr as written by a 64-bit
// little-endian build: twelve 8-byte fields (magic, version, r, s, skip,
// m, k_iv, blk, q, f0, has_A, nw), then u32 crc_h, crc_a, crc_hdr, padded
// to 112 bytes; crc_hdr covers the first 108 bytes with crc_hdr itself
// zero.  The m*nw words of the h-table follow (CRC crc_h), then the nw
// words of A when has_A (CRC crc_a).
constexpr u64 TSFC_MAGIC = 0x31434653544ULL;
constexpr size_t TSFC_HDR = 112;

struct Tsfc {
    u64 r = 0, s = 0, skip = 0, m = 0, k_iv = 0, blk = 0, q = 0, has_A = 0, nw = 0;
};

inline bool load_tsfc_one(const std::string &path, Tsfc &t) {
    std::vector<unsigned char> b;
    if (!read_file(path, b) || b.size() < TSFC_HDR) return false;
    auto u64at = [&](size_t off) {
        u64 v = 0;
        for (int i = 0; i < 8; i++) v |= (u64)b[off + i] << (8 * i);
        return v;
    };
    auto u32at = [&](size_t off) {
        uint32_t v = 0;
        for (int i = 0; i < 4; i++) v |= (uint32_t)b[off + i] << (8 * i);
        return v;
    };
    if (u64at(0) != TSFC_MAGIC) return false;
    std::vector<unsigned char> h(b.begin(), b.begin() + 108);
    memset(h.data() + 104, 0, 4);
    if (crc32(h.data(), h.size()) != u32at(104)) return false;
    t.r = u64at(16);
    t.s = u64at(24);
    t.skip = u64at(32);
    t.m = u64at(40);
    t.k_iv = u64at(48);
    t.blk = u64at(56);
    t.q = u64at(64);
    t.has_A = u64at(80);
    t.nw = u64at(88);
    if (t.nw != (t.r + 63) / 64 || t.m < 1 || t.m > 64 || t.k_iv < 1) return false;
    const size_t hb = (size_t)(t.m * t.nw * 8), ab = t.has_A ? (size_t)(t.nw * 8) : 0;
    if (b.size() < TSFC_HDR + hb + ab) return false;
    if (crc32(b.data() + TSFC_HDR, hb) != u32at(96)) return false;
    if (t.has_A && crc32(b.data() + TSFC_HDR + hb, ab) != u32at(100)) return false;
    return true;
}

// the newest valid generation of <base>, <base>.1, <base>.2
inline bool load_tsfc(const std::string &base, Tsfc &t, std::string *which = nullptr) {
    const char *sfx[3] = {"", ".1", ".2"};
    for (const char *x : sfx)
        if (load_tsfc_one(base + x, t)) {
            if (which) *which = base + x;
            return true;
        }
    return false;
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
    // s values that are not in the running cohort but must survive a
    // checkpoint: resumed s whose cohort has not run yet, and finished s
    // whose lines are not yet in the results file, at their last known
    // k_done
    std::map<u64, u64> carry;
    static double now() {
        using namespace std::chrono;
        return duration<double>(steady_clock::now().time_since_epoch()).count();
    }
    // gf2t_scan_params.progress
    static int cb(void *user, uint64_t k_done, const uint64_t *svals, size_t n) {
        Writer &W = *(Writer *)user;
        const bool hook = W.stop_after >= 0 && ++W.calls >= W.stop_after;
        const bool stop = stop_flag() || hook;
        for (size_t i = 0; i < n; i++) {
            auto it = W.carry.find(svals[i]);
            W.carry[svals[i]] = it == W.carry.end() ? k_done : std::max(it->second, k_done);
        }
        if (!W.path.empty() && (stop || now() - W.last >= W.every_s)) {
            Ckpt c;
            c.p = W.p;
            c.ent.assign(W.carry.begin(), W.carry.end());
            if (!save(W.path, c)) fprintf(stderr, "checkpoint %s: write failed\n", W.path.c_str());
            else if (W.verbose || stop)
                fprintf(stderr, "checkpoint %s: %zu s, running cohort of %zu clear through degree %llu\n",
                        W.path.c_str(), c.ent.size(), n, (unsigned long long)k_done);
            W.last = now();
        }
        if (stop) W.stopped = true;
        return stop ? 1 : 0;
    }
    // the lines of these s are in the results file: drop them
    void flushed(const std::vector<u64> &svals) {
        for (u64 s : svals) carry.erase(s);
    }
};

}  // namespace scan_ckpt
