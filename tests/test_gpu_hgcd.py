"""Device product service (gf2t_mul_host, gf2t_mat2_host) and the half-gcd
with device products, against the oracle's schoolbook products and
Euclid, up to the full r=136279841 size through the deep-scan path."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
P64 = C.POINTER(C.c_uint64)
R136 = 136279841


@pytest.fixture(scope="module")
def lib():
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import lib as L
    device.init(0)
    return L


def ptr(a):
    return a.ctypes.data_as(P64)


def rp(rng, bits, top=True):
    w = rng.integers(0, 2**64, O.words(bits), dtype=np.uint64)
    if bits & 63:
        w[-1] &= np.uint64((1 << (bits & 63)) - 1)
    if top:
        w[-1] |= np.uint64(1 << ((bits - 1) & 63))
    return w


@pytest.mark.parametrize("abits,bbits", [(1, 1), (64, 65), (1000, 3), (40000, 70001),
                                         (1 << 20, 333333)])
def test_mul_host_vs_oracle(lib, abits, bbits):
    rng = np.random.default_rng(abits ^ bbits)
    a, b = rp(rng, abits), rp(rng, bbits)
    out = np.zeros(lib.gf2t_mul_out_words(abits, bbits), dtype=np.uint64)
    assert lib.gf2t_mul_host(ptr(a), abits, ptr(b), bbits, ptr(out)) == 0
    assert np.array_equal(out, O.mul_comb(a, abits, b, bbits))


@pytest.fixture(params=[0, 1 << 30], ids=["copies", "zero-copy"])
def zc(lib, request):
    """the product service with its operands and results copied through
    device buffers (hm_zc_words 0) and read and written in pinned host
    memory directly (hm_zc_words above every size here)"""
    old = C.c_int64()
    assert lib.gf2t_tune_get(b"hm_zc_words", C.byref(old)) == 0
    assert lib.gf2t_tune(b"hm_zc_words", request.param) == 0
    yield request.param
    assert lib.gf2t_tune(b"hm_zc_words", old.value) == 0


@pytest.mark.parametrize("npairs,mb,vb", [(1, 300, 7000), (2, 40000, 100000), (3, 70001, 260000),
                                          (4, 64, 64)])
def test_mat2n_host_vs_products(lib, zc, npairs, mb, vb):
    """One matrix applied to npairs vector pairs of different lengths (one
    set of matrix transforms) equals the schoolbook products."""
    rng = np.random.default_rng(npairs * 1000 + mb)
    mbits = [mb - 3 * i if mb > 16 else mb for i in range(4)]
    m = [rp(rng, x) for x in mbits]
    vbits = [max(1, vb - 1000 * i) for i in range(2 * npairs)]
    v = [rp(rng, x) for x in vbits]
    ow = lib.gf2t_mul_out_words(max(mbits), max(vbits)) + 3
    outs = [np.full(ow, 0xDEAD, dtype=np.uint64) for _ in range(2 * npairs)]
    hm = (P64 * 4)(*[ptr(x) for x in m])
    hv = (P64 * (2 * npairs))(*[ptr(x) for x in v])
    ho = (P64 * (2 * npairs))(*[ptr(x) for x in outs])
    mb_arr = np.array(mbits, dtype=np.uint64)
    vb_arr = np.array(vbits, dtype=np.uint64)
    assert lib.gf2t_mat2n_host(hm, ptr(mb_arr), npairs, hv, ptr(vb_arr), ho, ow) == 0

    def prod(x, xb, y, yb):
        p = O.fft_mul(x, xb, y, yb)
        out = np.zeros(ow, dtype=np.uint64)
        out[:len(p)] = p
        return out

    for p in range(npairs):
        a, ab, b, bb = v[2 * p], vbits[2 * p], v[2 * p + 1], vbits[2 * p + 1]
        assert np.array_equal(outs[2 * p], prod(m[0], mbits[0], a, ab) ^ prod(m[1], mbits[1], b, bb))
        assert np.array_equal(outs[2 * p + 1],
                              prod(m[2], mbits[2], a, ab) ^ prod(m[3], mbits[3], b, bb))
    assert lib.gf2t_mat2n_host(hm, ptr(mb_arr), 5, hv, ptr(vb_arr), ho, ow) != 0


@pytest.mark.parametrize("mb,vb", [(100, 5000), (30000, 90000), (1, 1)])
def test_mat2_host_vs_products(lib, zc, mb, vb):
    rng = np.random.default_rng(mb + vb)
    m = [rp(rng, mb - i if mb > 4 else mb) for i in range(4)]
    mbits = np.array([mb - i if mb > 4 else mb for i in range(4)], dtype=np.uint64)
    a, b = rp(rng, vb), rp(rng, vb - 7 if vb > 8 else vb)
    bb = vb - 7 if vb > 8 else vb
    ow = lib.gf2t_mul_out_words(mb, vb)
    oa = np.zeros(ow, dtype=np.uint64)
    ob = np.zeros(ow, dtype=np.uint64)
    hm = (P64 * 4)(*[ptr(x) for x in m])
    assert lib.gf2t_mat2_host(hm, ptr(mbits), ptr(a), vb, ptr(b), bb, ptr(oa), ptr(ob), ow) == 0

    def prod(x, xb, y, yb):
        p = O.mul_comb(x, xb, y, yb)
        out = np.zeros(ow, dtype=np.uint64)
        out[:min(ow, len(p))] = p[:ow]
        return out
    wa = prod(m[0], int(mbits[0]), a, vb) ^ prod(m[1], int(mbits[1]), b, bb)
    wb = prod(m[2], int(mbits[2]), a, vb) ^ prod(m[3], int(mbits[3]), b, bb)
    assert np.array_equal(oa, wa) and np.array_equal(ob, wb)


def gcd_host(lib, a, b):
    cap = max(len(a), len(b)) + 1
    g = np.zeros(cap, dtype=np.uint64)
    d = lib.gf2t_gcd_host(ptr(a), len(a), ptr(b), len(b), ptr(g), cap)
    return d, g


@pytest.mark.parametrize("dev_bits", [0, 1 << 12])
def test_gcd_with_device_products_vs_oracle(lib, zc, dev_bits):
    assert lib.gf2t_tune(b"gcd_dev_bits", dev_bits) == 0
    try:
        rng = np.random.default_rng(dev_bits + 3)
        for trial in range(6):
            gb = int(rng.integers(1, 5000))
            g = rp(rng, gb)
            ab, bb = int(rng.integers(20000, 120000)), int(rng.integers(20000, 120000))
            a = O.mul_comb(g, gb, rp(rng, ab), ab)
            b = O.mul_comb(g, gb, rp(rng, bb), bb) if trial % 2 else rp(rng, bb)
            d1, g1 = O.gcd(a, b)
            d2, g2 = gcd_host(lib, a, b)
            n = O.words(d1 + 1)
            assert d1 == d2 and np.array_equal(g1[:n], g2[:n]), trial
    finally:
        lib.gf2t_tune(b"gcd_dev_bits", 1 << 18)


def test_full_size_gcd_finds_degree_41_factor(lib):
    # T = x^136279841 + x^474 + 1 has the least factor p3991ef16129 of
    # degree 41 (src/tsfactor_handoff.md:337-338) and no factor of degree
    # 1; gcd(T, x^(2^41) + x) is exactly that factor
    import mersenne_trinomial_amd as M
    x = O.x_residue(R136)
    f = M.square_fold(x, R136, 474, 41)
    f[0] ^= np.uint64(2)
    d, g = gcd_host(lib, O.trinomial(R136, 474), f)
    assert d == 41
    assert int(g[0]) == 0x3991ef16129 and not np.any(g[1:])
