"""The kernels' own arithmetic (gf2t_arith.hpp, shared verbatim by the
gfx950 kernels) built for the host and checked against the oracle on the
CPU, including every window-clipping edge.  No GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "mersenne_trinomial_amd", "csrc")
P = C.POINTER(C.c_uint64)


@pytest.fixture(scope="module")
def kl(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("kl") / "libkl.so")
    subprocess.check_call(["g++", "-O2", "-shared", "-fPIC", "-I" + CSRC, "-o", out,
                           os.path.join(HERE, "host", "kernel_logic.cpp")])
    lib = C.CDLL(out)
    lib.kl_square_fold.argtypes = [P, P, C.c_uint64, C.c_uint64]
    lib.kl_square_fold_fast.argtypes = [P, P, C.c_uint64, C.c_uint64]
    lib.kl_fold.argtypes = [P, C.c_uint64, P, C.c_uint64, C.c_uint64]
    lib.kl_square2_fold.argtypes = [P, P, C.c_uint64, C.c_uint64]
    lib.kl_square2_fold_units.argtypes = [P, P, C.c_uint64, C.c_uint64, P]
    lib.kl_square_fold_tiles.argtypes = [P, P, C.c_uint64, C.c_uint64, C.c_uint32]
    lib.kl_spread32p.restype = C.c_uint64
    lib.kl_spread32p.argtypes = [C.c_uint32]
    lib.kl_spread32.restype = C.c_uint64
    lib.kl_interleave32.restype = C.c_uint64
    lib.kl_interleave32.argtypes = [C.c_uint32, C.c_uint32]
    lib.kl_spread32.argtypes = [C.c_uint32]
    lib.kl_funnel64.restype = C.c_uint64
    lib.kl_funnel64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
    lib.kl_gf64_mul.restype = C.c_uint64
    lib.kl_gf64_mul.argtypes = [C.c_uint64, C.c_uint64]
    lib.kl_clmul32.restype = C.c_uint64
    lib.kl_clmul32.argtypes = [C.c_uint32, C.c_uint32]
    lib.kl_fmul.restype = C.c_uint32
    lib.kl_fmul.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_int]
    lib.kl_range_mask64.restype = C.c_uint64
    lib.kl_range_mask64.argtypes = [C.c_int, C.c_int]
    lib.kl_tile_of.restype = C.c_int64
    lib.kl_tile_of.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, P]
    lib.kl_gcd32.restype = C.c_uint32
    lib.kl_gcd32.argtypes = [C.c_uint32, C.c_uint32]
    lib.kl_inv32.restype = C.c_uint32
    lib.kl_inv32.argtypes = [C.c_uint32, C.c_uint32]
    lib.kl_mod_mersenne.restype = C.c_uint64
    lib.kl_mod_mersenne.argtypes = [C.c_uint64, C.c_int]
    lib.kl_minpoly_bm.restype = C.c_uint64
    lib.kl_minpoly_bm.argtypes = [C.c_uint32, C.c_uint64, C.c_int, C.c_uint32]
    lib.kl_count_classes.restype = C.c_uint64
    lib.kl_count_classes.argtypes = [C.c_int]
    lib.kl_count_classes_all.restype = C.c_uint64
    lib.kl_count_classes_all.argtypes = [C.c_int]
    for name, args in [("kl_fmul64", [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]),
                       ("kl_minpoly_bm64", [C.c_uint64, C.c_uint64, C.c_int, C.c_uint64]),
                       ("kl_mulmod64", [C.c_uint64, C.c_uint64, C.c_uint64]),
                       ("kl_inv64", [C.c_uint64, C.c_uint64]),
                       ("kl_gcd64", [C.c_uint64, C.c_uint64]),
                       ("kl_count_classes64", [C.c_int])]:
        getattr(lib, name).restype = C.c_uint64
        getattr(lib, name).argtypes = args
    return lib


def ptr(a):
    return a.ctypes.data_as(P)


def rand_res(rng, r):
    a = rng.integers(0, 2**64, O.words(r), dtype=np.uint64)
    if r & 63:
        a[-1] &= np.uint64((1 << (r & 63)) - 1)
    return a


def test_range_mask(kl):
    for lc in range(-130, 140, 3):
        for hc in range(-130, 140, 5):
            want = 0
            for b in range(64):
                if lc <= b < hc:
                    want |= 1 << b
            assert kl.kl_range_mask64(lc, hc) == want, (lc, hc)


def test_square_fold_all_s_small_r(kl):
    rng = np.random.default_rng(0)
    for r in (7, 31, 65, 127, 129, 131, 193, 257, 383):
        for s in range(1, (r - 1) // 2 + 1):
            a = rand_res(rng, r)
            out = np.zeros(O.words(r), np.uint64)
            kl.kl_square_fold(ptr(a), ptr(out), r, s)
            want = O.square_fold(a, r, s)
            assert np.array_equal(out, want), (r, s)
            kl.kl_square_fold_fast(ptr(a), ptr(out), r, s)
            assert np.array_equal(out, want), ("fast", r, s)


def test_square_fold_window_edges(kl):
    rng = np.random.default_rng(1)
    for r in (4423, 4097, 9689, 132049):
        for s in sorted({1, 2, 63, 64, 65, 127, 128, 129, 191, 192, 255, 256, 257, r // 3,
                         (r - 1) // 2 - 64, (r - 1) // 2 - 1, (r - 1) // 2}):
            a = rand_res(rng, r)
            out = np.zeros(O.words(r), np.uint64)
            kl.kl_square_fold(ptr(a), ptr(out), r, s)
            want = O.square_fold(a, r, s)
            assert np.array_equal(out, want), (r, s)
            kl.kl_square_fold_fast(ptr(a), ptr(out), r, s)
            assert np.array_equal(out, want), ("fast", r, s)


def test_square_fold_fast_full_size_sample(kl):
    rng = np.random.default_rng(9)
    r = 136279841
    for s in (474, 68139920, 1 << 20, 12345679):
        a = rand_res(rng, r)
        out = np.zeros(O.words(r), np.uint64)
        kl.kl_square_fold_fast(ptr(a), ptr(out), r, s)
        assert np.array_equal(out, O.square_fold(a, r, s)), s


def _spread_ref(x):
    return sum(((x >> k) & 1) << (2 * k) for k in range(32))


def test_spread_by_byte_permute(kl):
    rng = np.random.default_rng(11)
    xs = [0, 1, 0xFFFFFFFF, 0x80000000, 0x0F0F0F0F, 0xF0F0F0F0, 0x88888888, 0x77777777]
    xs += [int(v) for v in rng.integers(0, 2**32, 2000, dtype=np.uint64)]
    for x in xs:
        assert kl.kl_spread32p(x) == _spread_ref(x) == kl.kl_spread32(x), hex(x)
    for e, o in zip(xs, reversed(xs)):
        assert kl.kl_interleave32(e, o) == _spread_ref(e) | (_spread_ref(o) << 1), (hex(e), hex(o))


def test_funnel_shift(kl):
    rng = np.random.default_rng(12)
    for _ in range(500):
        a, b = (int(v) for v in rng.integers(0, 2**64, 2, dtype=np.uint64))
        for sh in range(64):
            want = ((a >> sh) | (b << (64 - sh))) & (2**64 - 1) if sh else a
            assert kl.kl_funnel64(a, b, sh) == want, (a, b, sh)


def test_two_squarings_per_tile(kl):
    """The two-squaring kernel's tiles (window plan, G phase into per-window
    copies, H phase from them; every read checked against what the G
    phase wrote) give (F^2)^2 mod T.  Small r puts the points s, 2s and r
    inside tiles at every offset; the single kernel's per-tile region path
    is checked at several tile sizes alongside."""
    rng = np.random.default_rng(13)
    cases = [(r, s) for r in (7, 31, 65, 127, 129, 131, 193, 257, 383)
             for s in range(1, (r - 1) // 2 + 1)]
    cases += [(r, s) for r in (4423, 4097, 9689, 132049, 300007)
              for s in sorted({1, 2, 63, 64, 65, 127, 128, 129, 255, 256, 257, r // 3,
                               (r - 1) // 2 - 64, (r - 1) // 2 - 1, (r - 1) // 2})]
    cases += [(300007, int(s)) for s in rng.integers(1, 150003, 40)]
    for r, s in cases:
        a = rand_res(rng, r)
        one = O.square_fold(a, r, s)
        want = O.square_fold(one, r, s)
        out = np.zeros(O.words(r), np.uint64)
        kl.kl_square2_fold(ptr(a), ptr(out), r, s)
        assert np.array_equal(out, want), (r, s)
        for tp in (1, 3, 512):
            kl.kl_square_fold_tiles(ptr(a), ptr(out), r, s, tp)
            assert np.array_equal(out, one), (r, s, tp)


def test_window_loads_every_r_mod_64(kl):
    """r = 4097..4160 covers every r mod 64.  Where r is not a multiple of
    64 the kernels read each window's three words as one 12-byte load,
    word q + 2 also when the shift is 0; the emulation's source aborts on
    any read past the 2 nw words, so this checks that those reads stay in
    bounds (and r = 4160 takes the other form), through the one-step
    tiles, the two-squaring tiles and the work units."""
    rng = np.random.default_rng(17)
    for r in range(4097, 4161):
        h = (r - 1) // 2
        for s in sorted({1, 2, 63, 64, 65, r // 3, h - 1, h, int(rng.integers(1, h + 1))}):
            a = rand_res(rng, r)
            one = O.square_fold(a, r, s)
            two = O.square_fold(one, r, s)
            out = np.zeros(O.words(r), np.uint64)
            kl.kl_square_fold_fast(ptr(a), ptr(out), r, s)
            assert np.array_equal(out, one), (r, s)
            for tp in (1, 512):
                kl.kl_square_fold_tiles(ptr(a), ptr(out), r, s, tp)
                assert np.array_equal(out, one), (r, s, tp)
            kl.kl_square2_fold(ptr(a), ptr(out), r, s)
            assert np.array_equal(out, two), (r, s)
            out2, _ = units(kl, a, r, s)
            assert np.array_equal(out2, two), (r, s)


def units(kl, a, r, s):
    out = np.zeros(O.words(r), np.uint64)
    st = np.zeros(3, np.uint64)
    kl.kl_square2_fold_units(ptr(a), ptr(out), r, s, ptr(st))
    return out, [int(x) for x in st]


def test_two_squarings_chained_units(kl):
    """The batched two-squaring kernel's work units: chains of tiles
    delta = off2 - off3 pairs apart, each tile taking window 3 (and window 5
    where the copy covers it) from the previous tile's copies.  Every read
    is checked against the G word the LDS word holds and every output pair
    must be formed once; s spans the unchained case (delta < 502), the
    switch at delta = 502 (s ~ 64256), one to many columns per period and
    chains cut into several units, at both parities of r, s and
    off4 - off5 - delta."""
    rng = np.random.default_rng(15)
    cases = [(r, s) for r in (7, 131, 383, 4423) for s in (1, (r - 1) // 3, (r - 1) // 2)]
    for r in (300007, 1000003, 1048573):
        h = (r - 1) // 2
        cases += [(r, s) for s in sorted({1, 2, 64127, 64128, 64255, 64256, 64257, 64383,
                                          64384, 65279, 65280, 65281, 65535, 65536, 130559,
                                          130560, 130561, r // 5, r // 4, r // 3, h - 1, h}
                                         & set(range(1, h + 1)))]
        cases += [(r, int(s)) for s in rng.integers(65000, h, 12)]
    chained = set()
    for r, s in cases:
        a = rand_res(rng, r)
        want = O.square_fold(O.square_fold(a, r, s), r, s)
        out, (nu, nt, formed) = units(kl, a, r, s)
        assert np.array_equal(out, want), (r, s)
        assert nu > 0 and nt >= nu and formed >= 2 * nt
        chained.add(s > 128 * 503)
    assert chained == {False, True}


def test_chained_units_form_fewer_windows(kl):
    """In [2s, r) a chained tile forms windows 1 and 2 only: at s ~ r/4
    the windows formed per tile fall from 3 + s/r (plain tiles) to about
    2 + s/r, up to the one column in about 30 whose tiles straddle s, 2s
    and r."""
    rng = np.random.default_rng(16)
    r = 8388593
    s = r // 4
    a = rand_res(rng, r)
    _, (nu, nt, formed) = units(kl, a, r, s)
    assert formed < 2.35 * nt, (nu, nt, formed)


def test_two_squarings_full_size_sample(kl):
    rng = np.random.default_rng(14)
    r = 136279841
    for s in (474, 68139920):
        a = rand_res(rng, r)
        out = np.zeros(O.words(r), np.uint64)
        kl.kl_square2_fold(ptr(a), ptr(out), r, s)
        want = O.square_fold(O.square_fold(a, r, s), r, s)
        assert np.array_equal(out, want), s
        out, _ = units(kl, a, r, s)
        assert np.array_equal(out, want), s
    a = rand_res(rng, r)
    out, _ = units(kl, a, r, 30000001)
    assert np.array_equal(out, O.square_fold(O.square_fold(a, r, 30000001), r, 30000001))


def test_fold_lengths(kl):
    rng = np.random.default_rng(2)
    for r, s in [(131, 65), (129, 64), (4423, 1200), (4097, 2048), (4097, 64), (4161, 127)]:
        for L in sorted({1, 63, 64, 65, r - 1, r, r + 1, r + s - 1, r + s, 2 * r - s,
                         2 * r - 2, 2 * r - 1}):
            v = rand_res(rng, L)
            out = np.zeros(O.words(r), np.uint64)
            kl.kl_fold(ptr(v), L, ptr(out), r, s)
            assert np.array_equal(out, O.reduce_bitserial(v, L, r, s)), (r, s, L)


def test_field_multiplies(kl):
    rng = np.random.default_rng(3)
    for a, b in rng.integers(0, 2**64, (3000, 2), dtype=np.uint64):
        assert kl.kl_gf64_mul(int(a), int(b)) == O.gf64_mul_bitserial(int(a), int(b))
        x, y = int(a) & 0xFFFFFFFF, int(b) & 0xFFFFFFFF
        assert kl.kl_clmul32(x, y) == O.lib().gf2o_clmul32(x, y)
    # dense operands: the column counts that the paired multiply-adds of
    # clmul32 rely on peak at all-ones (one dense operand and one sparse
    # one, then both dense with a few bits cleared)
    dense = [0xFFFFFFFF, 0xFFFFFFFE, 0x7FFFFFFF, 0xEFFFFFFF, 0xFFFF7FFF, 0x11111111,
             0x88888888, 0xF0F0F0F0, 0x0F0F0F0F, 0xAAAAAAAA, 0x55555555, 1, 0x80000000]
    for x in dense:
        for y in dense:
            assert kl.kl_clmul32(x, y) == O.lib().gf2o_clmul32(x, y), (hex(x), hex(y))
            xx, yy = (x << 32) | x, (y << 32) | (y ^ 0xFFFFFFFF)
            assert kl.kl_gf64_mul(xx, yy) == O.gf64_mul_bitserial(xx, yy), (hex(xx), hex(yy))
    for a, b in rng.integers(0, 2**32, (2000, 2), dtype=np.uint64):
        x, y = 0xFFFFFFFF ^ (1 << (int(a) & 31)), 0xFFFFFFFF & ~int(b) | 0xFFFF0FFF
        assert kl.kl_clmul32(x, y) == O.lib().gf2o_clmul32(x, y), (hex(x), hex(y))
    for d in (2, 5, 16, 17, 31, 32):
        q = O.find_primitive_poly(d)
        for a, b in rng.integers(0, 1 << d, (300, 2), dtype=np.uint64):
            assert kl.kl_fmul(int(a), int(b), q, d) == O.fmul(int(a), int(b), q, d)


def test_sieve_number_theory(kl):
    import math
    rng = np.random.default_rng(5)
    for d in (2, 3, 16, 31, 32):
        M = (1 << d) - 1
        for j in list(rng.integers(1, M, 300)) + [1, M - 1]:
            j = int(j)
            assert kl.kl_gcd32(j, M) == math.gcd(j, M)
            g = math.gcd(j, M)
            n = M // g
            if n > 1:
                assert kl.kl_inv32(j // g, n) == pow(j // g, -1, n)
            x = int(rng.integers(0, 2**63)) * 2 + 1
            assert kl.kl_mod_mersenne(x, d) == x % M


def _trace_mask(q, d):
    t = 0
    for i in range(d):
        y = 1 << i
        acc, cur = y, y
        for _ in range(d - 1):
            cur = O.fmul(cur, cur, q, d)
            acc ^= cur
        assert acc in (0, 1)
        t |= acc << i
    return t


def _minpoly_product(beta, q, d):
    # prod over the d conjugates (the reference's formulation)
    P = [1] + [0] * d
    root = beta
    for i in range(d):
        for k in range(i + 1, 0, -1):
            P[k] = P[k - 1] ^ O.fmul(root, P[k], q, d)
        P[0] = O.fmul(root, P[0], q, d)
        root = O.fmul(root, root, q, d)
    assert all(c in (0, 1) for c in P)
    return sum(c << k for k, c in enumerate(P))


def test_minpoly_bm_equals_product(kl):
    rng = np.random.default_rng(6)
    for d in (2, 3, 4, 5, 8, 13, 20, 31, 32):
        q = O.find_primitive_poly(d)
        tm = _trace_mask(q, d)
        M = (1 << d) - 1
        g = 2
        n = 0
        for j in rng.integers(1, M, 60):
            j = int(j)
            # skip subfield elements (orbit shorter than d)
            rot = [((j << i) | (j >> (d - i))) & M for i in range(1, d)]
            if j in rot:
                continue
            beta = 1
            e, b = j, 2
            while e:
                if e & 1:
                    beta = O.fmul(beta, b, q, d)
                b = O.fmul(b, b, q, d)
                e >>= 1
            want = _minpoly_product(beta, q, d)
            assert kl.kl_minpoly_bm(beta, q, d, tm) == want, (d, j)
            assert O.is_irreducible_small(want, d)
            n += 1
        assert n > 0


def test_tile_order_is_a_bijection(kl):
    nb = np.zeros(1, dtype=np.uint64)
    for T in (1, 7, 16, 17, 100, 4159, 8318):
        for delta in (0, 1, 255, 256, 257, 4000, 40000, 532343, 10**7):
            kl.kl_tile_of(T, delta, 256, 0, ptr(nb))
            tiles = [kl.kl_tile_of(T, delta, 256, x, ptr(nb)) for x in range(int(nb[0]) + 40)]
            got = sorted(t for t in tiles if t >= 0)
            assert got == list(range(T)), (T, delta)


def test_square_fold_fast_with_tile_order_covers_output(kl):
    # the kernels' squaring is per-pair; the order only permutes tiles.
    # Re-run the fast path in tile order on the host and compare.
    rng = np.random.default_rng(4)
    for r, s in [(132049, 66024), (132049, 1), (756839, 300000)]:
        a = rand_res(rng, r)
        want = O.square_fold(a, r, s)
        out = np.zeros(O.words(r), np.uint64)
        kl.kl_square_fold_fast(ptr(a), ptr(out), r, s)
        assert np.array_equal(out, want)


def test_class_enumeration_counts(kl):
    # the sieve tests only odd j below 2^(d-1); the class count must still
    # be the number of degree-d irreducibles (Moebius), as over all j
    for d in range(2, 25):
        assert kl.kl_count_classes(d) == O.count_irreducible(d), d
    for d in range(2, 19):
        assert kl.kl_count_classes_all(d) == O.count_irreducible(d), d


def _pmul(a, b):
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        b >>= 1
    return r


def _pmod(a, q):
    dq = q.bit_length() - 1
    while a and a.bit_length() - 1 >= dq:
        a ^= q << (a.bit_length() - 1 - dq)
    return a


def _irreducible(d, rng):
    while True:
        q = (1 << d) | int(rng.integers(0, 1 << 62)) % (1 << d) | 1
        if O.is_irreducible_small(q, d):
            return q


def test_wide_field_arithmetic(kl):
    # 32 < d <= 40: products against Python carry-less products, the
    # Berlekamp-Massey minimal polynomial is irreducible of degree d and
    # vanishes at beta, number theory against Python integers
    rng = np.random.default_rng(33)
    for d in (33, 34, 35, 36, 40):
        q = _irreducible(d, rng)
        tmask = 0
        for i in range(d):
            y, acc = 1 << i, 1 << i
            for _ in range(1, d):
                y = _pmod(_pmul(y, y), q)
                acc ^= y
            assert acc in (0, 1)
            tmask |= acc << i
        for _ in range(20):
            a = int(rng.integers(0, 1 << 62)) % (1 << d)
            b = int(rng.integers(0, 1 << 62)) % (1 << d)
            assert kl.kl_fmul64(a, b, q, d) == _pmod(_pmul(a, b), q)
        for _ in range(6):
            beta = int(rng.integers(2, 1 << 62)) % (1 << d)
            mask = kl.kl_minpoly_bm64(beta, q, d, tmask)
            if mask == 0:          # beta in a proper subfield (rare)
                continue
            assert mask.bit_length() - 1 == d and O.is_irreducible_small(mask, d)
            acc, pw = 0, 1
            for k in range(d + 1):
                if (mask >> k) & 1:
                    acc ^= pw
                pw = kl.kl_fmul64(pw, beta, q, d)
            assert acc == 0
    M = (1 << 34) - 1
    for _ in range(200):
        a, b = int(rng.integers(1, M)), int(rng.integers(1, M))
        assert kl.kl_mulmod64(a, b, M) == a * b % M
        assert kl.kl_gcd64(a, M) == np.gcd(a, M)
        if np.gcd(a, M) == 1:
            assert kl.kl_inv64(a, M) * a % M == 1
    for d in range(3, 23):
        assert kl.kl_count_classes64(d) == O.count_irreducible(d), d
