// host_poly.hpp -- GF(2)[x] arithmetic on the host CPU for the parts of
// the search that stay on the host in the reference too: interval gcds
// against T, least-degree resolution and equal-degree factorization of
// the (small) interval gcd, certificate hex (reference: src/gf2_gcd.h:
// 48-145, src/tsfactor.cu:932-987).  Polynomials are little-endian u64
// word vectors; the carry-less products use PCLMULQDQ.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

namespace gf2t_host {

typedef uint64_t u64;
typedef std::vector<u64> Poly;

int64_t deg(const Poly &a);
int64_t deg_w(const u64 *a, size_t n);
void trim(Poly &a);
bool is_zero(const Poly &a);
bool less(const Poly &a, const Poly &b);          // numeric order
bool equal(const Poly &a, const Poly &b);
void xor_shift(Poly &a, const Poly &b, u64 sh);   // a ^= b << sh (a grown as needed)
Poly mul(const Poly &a, const Poly &b);
Poly sqr(const Poly &a);
void mod_inplace(Poly &a, const Poly &g);         // a <- a mod g (deg g >= 0)
Poly divexact(const Poly &a, const Poly &g);      // a / g, remainder must be 0
Poly gcd(const Poly &a, const Poly &b);           // monic; gcd(0,0) = 0
Poly mulmod(const Poly &a, const Poly &b, const Poly &g);
Poly sqrmod(const Poly &a, const Poly &g);
std::string hex(const Poly &a);                   // lowercase, no leading zero
Poly from_words(const u64 *w, size_t n);
Poly trinomial(u64 r, u64 s);

// Reduction modulo a fixed g: Barrett (mu = x^(2n) div g, n = deg g)
// with Karatsuba / device products for large g, shift-and-add otherwise.
class Modulus {
public:
    explicit Modulus(const Poly &g);
    Poly reduce(const Poly &a) const;
    Poly sqr(const Poly &a) const;
    Poly mul(const Poly &a, const Poly &b) const;
    const Poly &poly() const { return g_; }
private:
    Poly g_, mu_;
    int64_t n_;
    bool barrett_;
};

// least factor degree d in (ka-1, kb] of g, whose irreducible factors all
// have degrees in [ka, kb], and the numerically least degree-d factor.
// h0 = x^(2^(ka-1)) mod g.  Degrees are tried up to deg g / 2 only (a
// least factor of a reducible g lies there; else g itself is the factor).
// Returns false if nothing is found.
bool resolve_least(const Poly &g, u64 ka, u64 kb, const Poly &h0, u64 &d, Poly &mask);

// numerically least irreducible factor of t, all of whose irreducible
// factors have degree exactly d (t squarefree)
Poly edf_least(const Poly &t, u64 d);

// Swan bound of the reference (src/tsfactor.cu:1101-1107)
u64 swan_rhigh(u64 r, u64 s, bool *swan);

// ---- subquadratic products and gcd (hgcd.cpp)
// a[...] ^= b[0, bw) << sh; a must hold bw + sh/64 + 1 words
void xor_shift_words(u64 *a, const u64 *b, size_t bw, u64 sh);
// Products where both operands have at least `bits` bits go to the
// device product service; 0 keeps every product on the host.
void set_device_product_bits(u64 bits);
u64 device_product_bits();
// The half-gcd's base-case size in bits (128..65536; below it the
// word-window rounds run) and Karatsuba's schoolbook base in words
// (8..512): gf2t_tune "gcd_base_bits" and "gcd_school_words"; false for
// a value out of range
bool set_gcd_base_bits(int64_t bits);
int64_t gcd_base_bits();
bool set_school_words(int64_t words);
int64_t school_words();
// whether chains take two squarings per launch ("sq_fuse"; square.hip)
bool square_fuse_enabled();
Poly mul_fast(const Poly &a, const Poly &b);       // Karatsuba / device FFT
// The 2x2 step matrix of a stretch of the remainder sequence: it maps
// (a, b) to (m00 a + m01 b, m10 a + m11 b).
struct HMat {
    Poly m00, m01, m10, m11;
};
// deg a > deg b >= -1: M and (c, d) = M (a, b), consecutive remainders
// of (a, b) with deg c >= ceil(deg a / 2) > deg d
// Host threads this process may use: the affinity mask, capped by
// GF2T_HOST_THREADS or else OMP_NUM_THREADS when set (the job's CPU share
// on a shared machine)
int host_threads();

// Deep-scan block size installed for r (gf2t_tune "block_m_<k>" covers
// 2^k <= r < 2^(k+1)); 0 = none
int tuned_block_m(u64 r);
bool set_tuned_block_m(int k, int m);

// Deep-scan cohorts split into this many staggered parts (gf2t_tune
// "scan_stagger"; 0 = automatic, -1 = never)
void set_scan_stagger(int parts);
int scan_stagger();

// want_M = false: the caller needs only (c, d), and the node's final
// matrix product M = S R2 is skipped (M is then unspecified)
void hgcd(const Poly &a, const Poly &b, HMat &M, Poly &c, Poly &d, bool want_M = true);
Poly gcd_fast(const Poly &a, const Poly &b);      // monic; same as gcd()

}  // namespace gf2t_host
