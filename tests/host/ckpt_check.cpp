// Round trip and corruption checks of the deep-scan checkpoint file
// (mersenne_trinomial_amd/cli/scan_ckpt.hpp).  Test code only.
#include "scan_ckpt.hpp"
#include <cstdio>
#include <cstdlib>

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string path = argv[1];
    scan_ckpt::Ckpt c;
    c.p.r = 136279841;
    c.p.skip = 32;
    c.p.m = 32;
    c.p.maxd = 2048;
    c.p.zq = 0;
    c.k_done = 1056;
    for (uint64_t s = 3; s < 40000; s += 7) c.s.push_back(s);
    if (!scan_ckpt::save(path, c)) return 3;
    scan_ckpt::Ckpt d;
    if (!scan_ckpt::load(path, d)) return 4;
    if (!(d.p == c.p) || d.k_done != c.k_done || d.s != c.s) return 5;
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
    puts("ckpt ok");
    return 0;
}
