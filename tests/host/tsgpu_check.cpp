// Drives the TsGPU shim of include/gf2t_engine.hpp through one blocked
// step of the reference's scan (src/tsfactor.cu:1242-1341) and dumps what
// each member produced, for tests/test_gpu_engine_shims.py to compare
// with the oracle.  Test code only.
//   tsgpu_check <r> <s> <m> <k> <outdir>
#include "gf2t_engine.hpp"
#include <cstdio>
#include <string>
#include <vector>

using gf2t::u64;

static void dump(const std::string &path, const std::vector<u64> &v) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f || fwrite(v.data(), 8, v.size(), f) != v.size()) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
}

static std::vector<u64> down(const u64 *d, size_t n) {
    std::vector<u64> v(n);
    gf2t::check(gf2t_memcpy_d2h(v.data(), d, n * 8, nullptr), "d2h");
    return v;
}

int main(int argc, char **argv) {
    if (argc != 6) return 2;
    const u64 r = strtoull(argv[1], 0, 10), s = strtoull(argv[2], 0, 10);
    const int m = atoi(argv[3]);
    const u64 k = strtoull(argv[4], 0, 10);
    const std::string out = argv[5];
    gf2t::check(gf2t_init(0), "gf2t_init");
    gf2t::TsGPU G;
    G.init(r);
    G.set_m(m);
    G.set_s(s);
    std::vector<u64> h0, h1, a1, a2, hb, ab;
    G.init_h_at(k);                      // block k .. k+m-1
    G.h_to_host(h0);
    dump(out + "/h0.bin", h0);
    G.horner_C();
    dump(out + "/c0.bin", down(G.dC, G.nw));
    G.reset_A();
    G.mul_A_by_C();
    G.A_to_host(a1);
    dump(out + "/a1.bin", a1);
    for (int t = 0; t < m; t++) G.square_all_step();   // block k+m .. k+2m-1
    G.h_to_host(h1);
    dump(out + "/h1.bin", h1);
    G.horner_C();
    dump(out + "/c1.bin", down(G.dC, G.nw));
    G.mul_A_by_C();
    G.A_to_host(a2);
    dump(out + "/a2.bin", a2);
    // full product and its fold, with the second block's C
    G.mul_full(G.dA, G.dC);
    dump(out + "/p.bin", down(G.dS0, gf2t::mul_out_words(r, r)));
    G.reduce_into(G.dTmp);
    dump(out + "/red.bin", down(G.dTmp, G.nw));
    G.square_slot(G.hslot(0), G.dTmp);
    dump(out + "/sq.bin", down(G.dTmp, G.nw));
    // host round trips and the hold copy
    G.h_from_host(h0);
    G.need_hold();
    G.copy_h(G.dHold, G.dH[G.hcur]);
    G.h_to_host(hb);
    if (hb != h0) { fprintf(stderr, "h round trip differs\n"); return 3; }
    std::vector<u64> hold = down(G.dHold, (size_t)m * G.hpitch);
    for (int j = 0; j < m; j++)
        for (size_t w = 0; w < G.nw; w++)
            if (hold[(size_t)j * G.hpitch + w] != h0[(size_t)j * G.nw + w]) {
                fprintf(stderr, "hold copy differs\n");
                return 4;
            }
    G.A_from_host(a1);
    G.A_to_host(ab);
    if (ab != a1) { fprintf(stderr, "A round trip differs\n"); return 5; }
    G.fini();
    puts("tsgpu ok");
    return 0;
}
