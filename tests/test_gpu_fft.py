"""(c) Cantor transform and products on the GPU against the oracle (and,
at full scale, residues modulo independent degree-61 irreducibles as in
the reference's bigtest).  Bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
R136 = 136279841


@pytest.fixture(scope="module")
def eng():
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    device.init(0)
    return M


def rand_bits(rng, bits):
    a = rng.integers(0, 2**64, max(1, O.words(bits)), dtype=np.uint64)
    if bits & 63:
        a[-1] &= np.uint64((1 << (bits & 63)) - 1)
    return a


@pytest.mark.parametrize("m", list(range(1, 21)))
def test_transform_bit_identical(eng, m):
    rng = np.random.default_rng(m)
    f = rng.integers(0, 2**64, 1 << m, dtype=np.uint64)
    F = eng.fft_fwd(f, m)
    assert np.array_equal(F, O.fft_fwd(f, m))
    assert np.array_equal(eng.fft_fwd(F, m, inverse=True), f)


@pytest.mark.parametrize("abits,bbits", [(1, 1), (31, 33), (64, 64), (65, 1), (1000, 2500),
                                         (4097, 3333), (1 << 16, (1 << 16) - 5),
                                         (100000, 70001), (1 << 20, 1 << 20),
                                         (3021377, 3021377)])
def test_product_vs_oracle(eng, abits, bbits):
    rng = np.random.default_rng(abits * 3 + bbits)
    a, b = rand_bits(rng, abits), rand_bits(rng, bbits)
    p = eng.mul(a, abits, b, bbits)
    want = O.fft_mul(a, abits, b, bbits) if abits > 5000 else O.mul_comb(a, abits, b, bbits)
    assert np.array_equal(p, want)


def _bits_for_chunks(rng, c):
    """a bit length whose 32-bit chunk count is c (last chunk partly used)"""
    return 32 * (c - 1) + int(rng.integers(1, 33))


# (2^mp, e, ca): products of N = ca + cb - 1 = 2^mp + e chunk coefficients,
# which run on 2^mp-point transforms plus a product of the top e chunks
# (wrapped products): e = 1, e = 2^mp / 4 (the largest wrapped), balanced
# and unbalanced operands, and e one past the limit (an ordinary product)
WRAPPED = [(11, 1, 1025), (12, 512, 2304), (12, 1024, 2560), (13, 2, 8000), (13, 3, 6000),
           (16, 3000, 34268), (17, 16384, 73728), (17, 32768, 81920), (15, 8193, 20480),
           (14, 700, 16384)]


@pytest.mark.parametrize("mp,e,ca", WRAPPED)
def test_wrapped_product_vs_oracle(eng, mp, e, ca):
    rng = np.random.default_rng(mp * 1000 + e)
    cb = (1 << mp) + e + 1 - ca
    abits, bbits = _bits_for_chunks(rng, ca), _bits_for_chunks(rng, cb)
    a, b = rand_bits(rng, abits), rand_bits(rng, bbits)
    print(f"2^{mp} + {e} coefficients: {abits} x {bbits} bits", flush=True)
    assert np.array_equal(eng.mul(a, abits, b, bbits), O.fft_mul(a, abits, b, bbits))
    assert np.array_equal(eng.mul(b, bbits, a, abits), O.fft_mul(a, abits, b, bbits))


@pytest.mark.parametrize("r", [530017, 589000, 655360, 700001])
def test_wrapped_slab_mulmod_vs_oracle(eng, r):
    """gf2t_mulmod_slab at r where the product's 2 ceil(r/32) - 1 chunk
    coefficients lie just above 2^15 (530017: 359 over, 589000: 4045 over,
    655360: 8191 over, the last wrapped) and, for comparison, past the
    wrapped range (700001), with F + x as the second operand on every
    other call."""
    import ctypes as C
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    n, nw = 3, O.words(r)
    pitch = (nw + 31) & ~31
    rng = np.random.default_rng(r)
    svals = np.array([1, 12345, (r - 1) // 2], np.uint64)
    a = [rand_bits(rng, r) for _ in range(n)]
    b = [rand_bits(rng, r) for _ in range(n)]

    def slab(vs):
        h = np.zeros(n * pitch, np.uint64)
        for i, v in enumerate(vs):
            h[i * pitch:i * pitch + nw] = v
        return DeviceBuffer.from_array(h)
    A, B, S = slab(a), slab(b), DeviceBuffer.from_array(svals)
    out = DeviceBuffer.words(n * pitch)
    work = DeviceBuffer(lib.gf2t_mulmod_slab_workspace_bytes(r, n))
    for add_x in (0, 1):
        check(lib.gf2t_mulmod_slab(A.ptr, pitch, B.ptr, pitch, add_x, r, S.ptr, out.ptr, pitch,
                                   n, work.ptr, None))
        got = out.download().reshape(n, pitch)
        for i in range(n):
            bb = b[i].copy()
            bb[0] ^= np.uint64(2 * add_x)
            want = O.mulmod(a[i], r, bb, r, r, int(svals[i]))
            assert np.array_equal(got[i, :nw], want), (add_x, i)


@pytest.mark.parametrize("r,s", [(4423, 1200), (132049, 1), (132049, 66024), (756839, 377)])
def test_mulmod_vs_oracle(eng, r, s):
    rng = np.random.default_rng(r ^ s)
    a, b = rand_bits(rng, r), rand_bits(rng, r)
    assert np.array_equal(eng.mulmod(a, r, b, r, r, s), O.mulmod(a, r, b, r, r, s))


def test_full_scale_product_residues(eng):
    # one 136279841 x 136279841-bit product (a wrapped product: 2^23-point
    # transforms plus a 2^18-point product of the top chunks), checked modulo
    # three degree-61 irreducibles (the reference's bigtest seeds,
    # src/cantor_ref.cpp:159-224), then folded mod x^r + x^474 + 1 and
    # compared with the oracle's linear fold
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    rng = np.random.default_rng(136)
    a, b = rand_bits(rng, R136), rand_bits(rng, R136)
    p = eng.mul(a, R136, b, R136)
    for seed in (0x1234567, 0xDEADBEEF, 0xC0FFEE):
        q = O.find_irreducible(61, seed)
        ra, rb = O.residue(a, R136, q, 61), O.residue(b, R136, q, 61)
        prod = O.mul_comb(np.array([ra], np.uint64), 64, np.array([rb], np.uint64), 64)
        assert O.residue(p, 2 * R136, q, 61) == O.residue(prod, 128, q, 61), hex(seed)
    src = DeviceBuffer.from_array(p)
    out = DeviceBuffer.words(O.words(R136))
    check(lib.gf2t_fold(src.ptr, 2 * R136 - 1, out.ptr, R136, 474, None))
    assert np.array_equal(out.download(), O.reduce_fold(p, 2 * R136 - 1, R136, 474))


def test_square_via_fft_equals_spread_square(eng):
    # the reference's bench cross-check (src/trinomial_stage2.cu:165-174)
    r, s = 756839, 12345
    rng = np.random.default_rng(1)
    a = rand_bits(rng, r)
    assert np.array_equal(eng.mulmod(a, r, a, r, r, s), eng.square_fold(a, r, s, 1))


def test_trin_gpu_interval_product(eng):
    # ACC = prod_{k=5..12} (F_k + x) at r=4423 s=1200
    # (src/trinomial_stage2.cu:109-137), against the oracle
    r, s = 4423, 1200
    G = eng.TrinGPU()
    G.init(r)
    G.set_s(s)
    G.set_x()
    G.reset_acc()
    F = O.x_residue(r)
    acc = np.zeros(O.words(r), np.uint64)
    acc[0] = 1
    for k in range(1, 13):
        G.square_A()
        F = O.square_fold(F, r, s)
        if k >= 5:
            G.accumulate()
            t = F.copy()
            t[0] ^= np.uint64(2)
            acc = O.mulmod(acc, r, t, r, r, s)
    assert np.array_equal(G.A(), F)
    assert np.array_equal(G.ACC(), acc)
    G.fini()


@pytest.mark.parametrize("m", [6, 9, 12, 14, 15, 17])
def test_every_tile_size_bit_identical(eng, m):
    """The transform is the same whatever tile the passes use: every tile
    size from 2^6 to 2^13 elements (gf2t_tune "fft_tile_<m>"), with 2, 8
    and 16 elements per thread (gf2t_tune "fft_ept") and 4, 5 and 6
    contiguous low index bits per tile (gf2t_tune "fft_low_bits"), forward
    against the oracle and back, and a product through the same plans."""
    from mersenne_trinomial_amd._capi import lib
    import ctypes as C
    rng = np.random.default_rng(100 + m)
    f = rng.integers(0, 2**64, 1 << m, dtype=np.uint64)
    want = O.fft_fwd(f, m)
    bits = (1 << (m + 4)) - 77         # a product that needs 2^m points
    a, b = rand_bits(rng, bits), rand_bits(rng, 29)
    pwant = O.fft_mul(a, bits, b, 29)
    old, old_ept, old_low = C.c_int64(), C.c_int64(), C.c_int64()
    assert lib.gf2t_tune_get(f"fft_tile_{m}".encode(), C.byref(old)) == 0
    assert lib.gf2t_tune_get(b"fft_ept", C.byref(old_ept)) == 0
    assert lib.gf2t_tune_get(b"fft_low_bits", C.byref(old_low)) == 0
    try:
        for ept, low in ((2, 4), (8, 5), (16, 6)):
            assert lib.gf2t_tune(b"fft_ept", ept) == 0
            assert lib.gf2t_tune(b"fft_low_bits", low) == 0
            for T in range(6, 14):
                print(f"m={m} tile 2^{T}, {ept} elements per thread, {low} low bits", flush=True)
                assert lib.gf2t_tune(f"fft_tile_{m}".encode(), T) == 0
                F = eng.fft_fwd(f, m)
                assert np.array_equal(F, want), (T, ept, low)
                assert np.array_equal(eng.fft_fwd(F, m, inverse=True), f), (T, ept, low)
                assert np.array_equal(eng.mul(a, bits, b, 29), pwant), (T, ept, low)
    finally:
        assert lib.gf2t_tune(f"fft_tile_{m}".encode(), old.value) == 0
        assert lib.gf2t_tune(b"fft_ept", old_ept.value) == 0
        assert lib.gf2t_tune(b"fft_low_bits", old_low.value) == 0
    assert lib.gf2t_tune(f"fft_tile_{m}".encode(), 5) != 0
    assert lib.gf2t_tune(b"fft_ept", 3) != 0
    assert lib.gf2t_tune(b"fft_low_bits", 3) != 0 and lib.gf2t_tune(b"fft_low_bits", 7) != 0
    assert lib.gf2t_tune(f"fft_tile_{m}".encode(), 14) != 0
