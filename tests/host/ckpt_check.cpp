// Round trip and corruption checks of the deep-scan checkpoint file
// (mersenne_trinomial_amd/cli/scan_ckpt.hpp), and the readers of the
// results file and of the reference's TSFC checkpoint.  Test code only.
//   ckpt_check <path>            round trip + corruption
//   ckpt_check tsfc <base>       print the TSFC header fields load_tsfc reads
//   ckpt_check done <results>    print the s values read_done takes
//   ckpt_check load <path>       print the (s, k_done) pairs load returns
#include "scan_ckpt.hpp"
#include <cinttypes>
#include <cstdio>
#include <cstdlib>

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    if (argc == 3 && std::string(argv[1]) == "tsfc") {
        scan_ckpt::Tsfc t;
        std::string which;
        if (!scan_ckpt::load_tsfc(argv[2], t, &which)) {
            puts("tsfc invalid");
            return 0;
        }
        printf("tsfc %s r=%" PRIu64 " s=%" PRIu64 " skip=%" PRIu64 " m=%" PRIu64 " k_iv=%" PRIu64
               " blk=%" PRIu64 " q=%" PRIu64 " has_A=%" PRIu64 " nw=%" PRIu64 "\n",
               which.c_str(), t.r, t.s, t.skip, t.m, t.k_iv, t.blk, t.q, t.has_A, t.nw);
        return 0;
    }
    if (argc == 3 && std::string(argv[1]) == "done") {
        std::set<uint64_t> done;
        scan_ckpt::read_done(argv[2], done);
        printf("done");
        for (uint64_t s : done) printf(" %" PRIu64, s);
        printf("\n");
        return 0;
    }
    if (argc == 3 && std::string(argv[1]) == "load") {
        scan_ckpt::Ckpt d;
        if (!scan_ckpt::load(argv[2], d)) {
            puts("load failed");
            return 0;
        }
        printf("load r=%" PRIu64 " skip=%" PRIu64 " m=%" PRIu64 " maxd=%" PRIu64 " zq=%" PRIu64,
               d.p.r, d.p.skip, d.p.m, d.p.maxd, d.p.zq);
        for (const auto &e : d.ent) printf(" %" PRIu64 ":%" PRIu64, e.first, e.second);
        printf("\n");
        return 0;
    }
    const std::string path = argv[1];
    scan_ckpt::Ckpt c;
    c.p.r = 136279841;
    c.p.skip = 32;
    c.p.m = 32;
    c.p.maxd = 2048;
    c.p.zq = 0;
    for (uint64_t s = 3; s < 40000; s += 7) c.ent.emplace_back(s, 1056 + (s % 5) * 32);
    if (!scan_ckpt::save(path, c)) return 3;
    scan_ckpt::Ckpt d;
    if (!scan_ckpt::load(path, d)) return 4;
    if (!(d.p == c.p) || d.ent != c.ent) return 5;
    // standard CRC-32 check value
    if (scan_ckpt::crc32("123456789", 9) != 0xCBF43926u) return 6;
    // flip one bit in the body: the CRC must reject it
    FILE *f = fopen(path.c_str(), "r+b");
    if (!f) return 7;
    fseek(f, 60, SEEK_SET);
    int ch = fgetc(f);
    fseek(f, 60, SEEK_SET);
    fputc(ch ^ 0x10, f);
    fclose(f);
    if (scan_ckpt::load(path, d)) return 8;
    // truncated file
    if (!scan_ckpt::save(path, c)) return 9;
    f = fopen(path.c_str(), "r+b");
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fclose(f);
    if (truncate(path.c_str(), n - 9) != 0) return 10;
    if (scan_ckpt::load(path, d)) return 11;
    // the writer keeps resumed s that have not run, and drops s once
    // their lines are flushed
    scan_ckpt::Writer W;
    W.path = path;
    W.p = c.p;
    W.every_s = 0;   // every call writes
    W.carry = {{5, 100}, {9, 200}};
    const uint64_t cohort[2] = {11, 13};
    scan_ckpt::Writer::cb(&W, 64, cohort, 2);
    if (!scan_ckpt::load(path, d)) return 12;
    const std::vector<std::pair<uint64_t, uint64_t>> want = {{5, 100}, {9, 200}, {11, 64}, {13, 64}};
    if (d.ent != want) return 13;
    W.flushed({5, 11});
    scan_ckpt::Writer::cb(&W, 128, cohort + 1, 1);
    if (!scan_ckpt::load(path, d)) return 14;
    const std::vector<std::pair<uint64_t, uint64_t>> want2 = {{9, 200}, {13, 128}};
    if (d.ent != want2) return 15;
    puts("ckpt ok");
    return 0;
}
