"""Host-side GF(2)[x] pieces of the deep scan (gcd, equal-degree
factorization, Swan bound) and the certificate formats, against the
oracle.  No GPU: these run on the host CPU through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

P = C.POINTER(C.c_uint64)


@pytest.fixture(scope="module")
def lib():
    from mersenne_trinomial_amd._capi import lib
    return lib


def ptr(a):
    return a.ctypes.data_as(P)


def rand_poly(rng, bits):
    a = rng.integers(0, 2**64, O.words(bits), dtype=np.uint64)
    if bits & 63:
        a[-1] &= np.uint64((1 << (bits & 63)) - 1)
    a[-1] |= np.uint64(1 << ((bits - 1) & 63))
    return a


def test_gcd_host_vs_oracle(lib):
    rng = np.random.default_rng(1)
    for _ in range(30):
        g = rand_poly(rng, int(rng.integers(2, 300)))
        a = O.mul_comb(g, len(g) * 64, rand_poly(rng, 2000), 2048)
        b = O.mul_comb(g, len(g) * 64, rand_poly(rng, 1500), 1536)
        cap = max(len(a), len(b)) + 1
        out = np.zeros(cap, dtype=np.uint64)
        d = lib.gf2t_gcd_host(ptr(a), len(a), ptr(b), len(b), ptr(out), cap)
        d2, g2 = O.gcd(a, b)
        assert d == d2
        assert np.array_equal(out[: len(g2)], g2)
    # a full-size trinomial gcd at r = 132049 against a random residue
    r, s = 132049, 1201
    t = O.trinomial(r, s)
    a = rand_poly(rng, r - 5)
    out = np.zeros(len(t) + 1, dtype=np.uint64)
    d = lib.gf2t_gcd_host(ptr(a), len(a), ptr(t), len(t), ptr(out), len(out))
    assert d == O.gcd(a, t)[0]


@pytest.mark.parametrize("seed", range(4))
def test_half_gcd_vs_euclid(lib, seed):
    # sizes that cross the half-gcd base case several times over, planted
    # common factors of every size, coprime pairs, equal degrees
    rng = np.random.default_rng(100 + seed)
    for trial in range(40):
        gb = int(rng.integers(1, 6000))
        g = rand_poly(rng, gb)
        ab, bb = int(rng.integers(1, 30000)), int(rng.integers(1, 30000))
        if trial % 5 == 4:
            bb = ab
        a = O.mul_comb(g, gb, rand_poly(rng, ab), ab)
        b = O.mul_comb(g, gb, rand_poly(rng, bb), bb) if trial % 3 else rand_poly(rng, bb)
        cap = max(len(a), len(b)) + 1
        out = np.zeros(cap, dtype=np.uint64)
        d = lib.gf2t_gcd_host(ptr(a), len(a), ptr(b), len(b), ptr(out), cap)
        d2, g2 = O.gcd(a, b)
        n = O.words(d2 + 1)
        assert d == d2, (seed, trial)
        assert np.array_equal(out[:n], g2[:n]) and not np.any(out[n:])


@pytest.mark.parametrize("base,school", [(128, 8), (1024, 48), (4096, 96), (16384, 512)])
def test_half_gcd_every_base_size(lib, base, school):
    """The gcd is the same whatever base case and schoolbook sizes the
    half-gcd uses (gf2t_tune "gcd_base_bits" from 128 to 16384 bits,
    "gcd_school_words" from 8 to 512 words): planted common factors and
    coprime pairs across the base case, and a trinomial gcd at r = 132049."""
    old_b, old_s = C.c_int64(), C.c_int64()
    assert lib.gf2t_tune_get(b"gcd_base_bits", C.byref(old_b)) == 0
    assert lib.gf2t_tune_get(b"gcd_school_words", C.byref(old_s)) == 0
    rng = np.random.default_rng(base + school)
    try:
        assert lib.gf2t_tune(b"gcd_base_bits", base) == 0
        assert lib.gf2t_tune(b"gcd_school_words", school) == 0
        for trial in range(12):
            gb = int(rng.integers(1, 3000))
            ab, bb = int(rng.integers(1, 40000)), int(rng.integers(1, 40000))
            g = rand_poly(rng, gb)
            a = O.mul_comb(g, gb, rand_poly(rng, ab), ab)
            b = O.mul_comb(g, gb, rand_poly(rng, bb), bb) if trial % 3 else rand_poly(rng, bb)
            cap = max(len(a), len(b)) + 1
            out = np.zeros(cap, dtype=np.uint64)
            d = lib.gf2t_gcd_host(ptr(a), len(a), ptr(b), len(b), ptr(out), cap)
            d2, g2 = O.gcd(a, b)
            n = O.words(d2 + 1)
            assert d == d2, trial
            assert np.array_equal(out[:n], g2[:n]) and not np.any(out[n:])
        r, s = 132049, 1201
        t = O.trinomial(r, s)
        a = rand_poly(rng, r - 5)
        out = np.zeros(len(t) + 1, dtype=np.uint64)
        assert lib.gf2t_gcd_host(ptr(a), len(a), ptr(t), len(t), ptr(out), len(out)) == \
            O.gcd(a, t)[0]
    finally:
        assert lib.gf2t_tune(b"gcd_base_bits", old_b.value) == 0
        assert lib.gf2t_tune(b"gcd_school_words", old_s.value) == 0
    for bad in (127, 65537):
        assert lib.gf2t_tune(b"gcd_base_bits", bad) != 0
    for bad in (7, 513):
        assert lib.gf2t_tune(b"gcd_school_words", bad) != 0


def test_half_gcd_trinomial_planted_factor(lib):
    # T = x^756839 + x^s + 1 against (residue with a planted small factor
    # of T): the gcd is that factor
    rng = np.random.default_rng(9)
    r = 756839
    best, _ = O.sieve(r, 14, 2000)
    s = int(np.nonzero(best[1:] != O.NONE)[0][-1]) + 1
    f = O.poly_from_hex(f"{int(best[s]) & 0xFFFFFFFFFFFF:x}")
    d = int(best[s]) >> 48
    t = O.trinomial(r, s)
    b = O.mul_comb(f, d + 1, rand_poly(rng, r - 100), r - 100)
    out = np.zeros(len(t) + 1, dtype=np.uint64)
    got = lib.gf2t_gcd_host(ptr(t), len(t), ptr(b), len(b), ptr(out), len(out))
    want, gw = O.gcd(t, b)
    assert got == want and want >= d
    assert np.array_equal(out[:len(gw)], gw)


def irreducibles(d, count, rng):
    found = []
    while len(found) < count:
        m = int(rng.integers(1 << d, 2 << d)) | 1
        if O.is_irreducible_small(m, d) and m not in found:
            found.append(m)
    return found


@pytest.mark.parametrize("d,k", [(5, 2), (13, 3), (24, 2), (31, 2), (40, 3), (61, 2)])
def test_edf_least(lib, d, k):
    rng = np.random.default_rng(d * 10 + k)
    fs = irreducibles(d, k, rng)
    t = np.array([1], dtype=np.uint64)
    for f in fs:
        t = O.mul_comb(t, len(t) * 64, np.array([f], dtype=np.uint64), 64)
        t = t[: O.words(O.deg(t) + 1)]
    out = np.zeros(O.words(d + 1) + 1, dtype=np.uint64)
    assert lib.gf2t_edf_least(ptr(t), len(t), d, ptr(out), len(out)) == 0
    assert int(out[0]) == min(fs)
    assert int(O.edf_least(t, d)[0]) == min(fs)


def test_swan_bound_values():
    import mersenne_trinomial_amd as M
    assert M.swan_bound(132049, 2) == (66024, False)      # s even, s | 2r
    assert M.swan_bound(132049, 5)[0] == 132049 // 3
    assert M.swan_bound(132049, 4)[0] == 132049 // 3      # 4 does not divide 2r
    assert M.swan_bound(4423, 1)[0] == 4423 // 3          # 4423 = 7 (mod 8), (r-s) does not divide 2r
    assert M.swan_bound(4229, 10)[0] == 4229 // 2         # 4229 = 5 (mod 8): no Swan bound


def test_cert_formats(tmp_path):
    from mersenne_trinomial_amd import certs
    best = np.full(11, certs.NONE, dtype=np.uint64)
    best[3] = (2 << 48) | 7
    best[7] = (5 << 48) | 0x29
    best[9] = 0
    lines = certs.sieve_lines(best)
    assert lines == ["3 2 p7", "7 5 p29"]
    out = certs.write_sieve_outputs(str(tmp_path / "x"), best, found_bin=True)
    assert open(out["cert"]).read().split("\n")[:2] == lines
    assert list(certs.read_survivors(out["survivors"])) == [1, 2, 4, 5, 6, 8, 10]
    back = certs.read_found_bin(out["found_bin"], 10)
    assert np.array_equal(back[1:], best[1:])
    assert certs.parse_line("474 41 p3991ef16129") == (474, "p", 41, "3991ef16129")
    assert certs.parse_line("12 u") == (12, "u", None, None)
    seeded = certs.best_from_survivors([2, 5], 6, 30)
    assert seeded[2] == certs.NONE and seeded[5] == certs.NONE and seeded[3] == np.uint64(29 << 48)


def test_host_products_vs_bitserial(tmp_path):
    """The gcd's host products (the 512-bit output-block schoolbook and the
    128-bit column one, Karatsuba with its fused recombination, unbalanced
    blocks, parallel halves) against a bit-serial product, with sizes around
    every switch point and eight-word block edge, zero words and all-ones
    words."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    csrc = os.path.join(root, "mersenne_trinomial_amd", "csrc")
    exe = tmp_path / "host_mul_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-mpclmul", "-msse4.1", "-pthread",
                           "-I" + os.path.join(root, "include"), "-I" + csrc, "-o", str(exe),
                           os.path.join(root, "tests", "host", "host_mul_check.cpp"),
                           os.path.join(csrc, "hgcd.cpp"), os.path.join(csrc, "host_poly.cpp")])
    # the 512-bit schoolbook where the CPU has AVX-512 VPCLMULQDQ, and the 128-bit one
    for env in ({}, {"GF2T_NO_VPCLMUL": "1"}):
        p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120,
                           env={**os.environ, **env})
        assert p.returncode == 0 and "host products ok" in p.stdout, (env, p.stdout + p.stderr)
