"""The oracle against the reference's own code, built from its sources in
oracle/_ref (skipped where the reference tree was not available to build
it).  No GPU."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built")


def test_basis_equal():
    rb = np.zeros(64, dtype=np.uint64)
    assert O.ref().ref_cantor_basis(O._ptr(rb)) == 64
    assert np.array_equal(rb, O.cantor_basis())


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 12, 16])
def test_transform_equal(m):
    rng = np.random.default_rng(100 + m)
    f = rng.integers(0, 2**64, 1 << m, dtype=np.uint64)
    a = O.fft_fwd(f, m)
    b = f.copy()
    O.ref().ref_fft_fwd(O._ptr(b), m)
    assert np.array_equal(a, b)
    c = a.copy()
    O.ref().ref_fft_inv(O._ptr(c), m)
    assert np.array_equal(c, f)


@pytest.mark.parametrize("abits,bbits", [(1, 1), (31, 33), (1000, 2500), (100000, 70001),
                                         (1 << 20, (1 << 20) - 17)])
def test_product_equal(abits, bbits):
    rng = np.random.default_rng(abits)
    a = rng.integers(0, 2**64, O.words(abits), dtype=np.uint64)
    b = rng.integers(0, 2**64, O.words(bbits), dtype=np.uint64)
    if abits & 63:
        a[-1] &= np.uint64((1 << (abits & 63)) - 1)
    if bbits & 63:
        b[-1] &= np.uint64((1 << (bbits & 63)) - 1)
    p = O.fft_mul(a, abits, b, bbits)
    q = np.zeros_like(p)
    O.ref().ref_fft_mul(O._ptr(a), abits, O._ptr(b), bbits, O._ptr(q))
    assert np.array_equal(p, q)


def test_gf64_equal():
    rng = np.random.default_rng(1)
    for a, b in rng.integers(0, 2**64, (500, 2), dtype=np.uint64):
        assert O.ref().ref_gf64_mul(int(a), int(b)) == O.gf64_mul(int(a), int(b))


def test_gcd_equal():
    rng = np.random.default_rng(4)
    for _ in range(10):
        g = rng.integers(0, 2**64, 2, dtype=np.uint64)
        a = O.mul_comb(g, 128, rng.integers(0, 2**64, 9, dtype=np.uint64), 576)
        b = O.mul_comb(g, 128, rng.integers(0, 2**64, 7, dtype=np.uint64), 448)
        d1, g1 = O.gcd(a, b)
        g2 = np.zeros(len(a) + 1, dtype=np.uint64)
        d2 = O.ref().ref_gcd_naive(O._ptr(a), len(a), O._ptr(b), len(b), O._ptr(g2), len(g2))
        assert d1 == d2
        assert np.array_equal(g1, g2[: len(g1)])


@pytest.mark.parametrize("r,depth,smax", [(101, 9, 100), (132049, 16, 66024),
                                          (3021377, 18, 200000)])
def test_sieve_lines_equal(r, depth, smax):
    best, _ = O.sieve(r, depth, smax, nthreads=4)
    assert O.cert_lines(best) == O.ref_sieve_lines(r, depth, smax)
