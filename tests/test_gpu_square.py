"""(b) squaring+fold on the GPU against the oracle, the known answers and
the reference's golden residue.  Bit-exact."""
import hashlib

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


def rand_res(rng, r):
    a = rng.integers(0, 2**64, O.words(r), dtype=np.uint64)
    if r & 63:
        a[-1] &= np.uint64((1 << (r & 63)) - 1)
    return a


# the reference's squaring selftest pairs (src/tsfactor.cu:1423-1480)
SQ_CASES = [(r, s) for r in (4423, 4159, 4097, 4229, 9689)
            for s in (1, 2, 3, 63, 64, 65, 129, r // 3, (r - 1) // 2 - 1, (r - 1) // 2)]
SQ_CASES += [(31, 3), (31, 15), (67, 33), (129, 64), (131, 1), (191, 95), (193, 96)]


@pytest.mark.parametrize("r,s", SQ_CASES)
def test_square_one_step(eng, r, s):
    rng = np.random.default_rng(r * 7919 + s)
    a = rand_res(rng, r)
    assert np.array_equal(eng.square_fold(a, r, s, 1), O.square_fold(a, r, s))


def test_frobenius_r31_known_answer(eng):
    x = O.x_residue(31)
    for k in range(1, 41):
        got = eng.square_fold(x, 31, 3, k)
        assert np.array_equal(got, O.square_fold_steps(x, 31, 3, k))
    assert list(eng.square_fold(x, 31, 3, 31)) == [2]


@pytest.mark.parametrize("r,s,k", [(132049, 1, 50), (132049, 66024, 33), (756839, 3000, 20),
                                   (3021377, 1510688, 12)])
def test_chain_vs_oracle(eng, r, s, k):
    x = O.x_residue(r)
    assert np.array_equal(eng.square_fold(x, r, s, k), O.square_fold_steps(x, r, s, k))


def test_full_size_golden_residue(eng):
    # F_32 = x^(2^32) mod x^136279841 + x^474 + 1 from the reference's
    # stage-2 driver: md5 of the 17,034,984-byte residue
    f = eng.square_fold(O.x_residue(R136), R136, 474, 32)
    assert hashlib.md5(f.tobytes()).hexdigest() == "ff5094d371880811cb34b354da3af4a7"


def test_full_size_random_s(eng):
    # two steps: one launch of the two-squaring kernel
    rng = np.random.default_rng(2024)
    for s in (1, 2, 68139919, 68139920, int(rng.integers(3, 68139920))):
        a = rand_res(rng, R136)
        assert np.array_equal(eng.square_fold(a, R136, s, 2), O.square_fold_steps(a, R136, s, 2)), s


@pytest.mark.parametrize("r", [4423, 9689, 132049])
def test_two_squarings_every_s(eng, r):
    """Every s of small r through one two-squaring launch, so that tiles
    meet s, 2s and r at every offset."""
    rng = np.random.default_rng(r)
    step = {4423: 1, 9689: 7}.get(r, 997)
    for s in range(1, (r - 1) // 2 + 1, step):
        a = rand_res(rng, r)
        assert np.array_equal(eng.square_fold(a, r, s, 2), O.square_fold_steps(a, r, s, 2)), s


@pytest.mark.parametrize("r", [1048573, 2000003])
def test_two_squarings_large_s(eng, r):
    """One two-squaring launch for s from about 128 tiles' worth of pairs
    (s ~ 64256, where the window offsets off2 - off3 pass one chained tile
    width) up to (r - 1)/2, at both parities, against the oracle."""
    rng = np.random.default_rng(r + 1)
    h = (r - 1) // 2
    svals = {64127, 64128, 64255, 64256, 64257, 64383, 64384, 64900, 65024, 65152, 65280,
             130000, h - 1, h}
    svals |= {int(x) for x in rng.integers(65000, h, 40)}
    for s in sorted(svals):
        a = rand_res(rng, r)
        assert np.array_equal(eng.square_fold(a, r, s, 2), O.square_fold_steps(a, r, s, 2)), s


@pytest.fixture(params=[(1, 0, 0, 0), (2, 0, 0, 0), (4, 0, 1, 0), (1, 1, 0, 0), (2, 1, 1, 0),
                        (4, 1, 0, 0), (2, 0, 1, 1), (2, 0, 0, 1), (1, 1, 1, 1)],
                ids=lambda p: f"ppt{p[0]}-order{p[1]}-nt{p[2]}-fuse{p[3]}")
def tuned(request, eng):
    """Single-squaring launch shapes (fuse0) and the two-squaring kernel
    (fuse1, with the single kernel for odd remainders)."""
    from mersenne_trinomial_amd._capi import check, lib
    ppt, order, nt, fuse = request.param
    check(lib.gf2t_tune(b"sq_ppt", ppt))
    check(lib.gf2t_tune(b"sq_order", order))
    check(lib.gf2t_tune(b"sq_nt", nt))
    check(lib.gf2t_tune(b"sq_fuse", fuse))
    yield request.param
    check(lib.gf2t_tune(b"sq_ppt", 2))
    check(lib.gf2t_tune(b"sq_order", 0))
    check(lib.gf2t_tune(b"sq_nt", 1))
    check(lib.gf2t_tune(b"sq_fuse", 1))


def test_batch_uneven_frontiers_mixed_r(eng, tuned):
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    import ctypes as C
    rng = np.random.default_rng(77)
    cases = [(4423, 1200, 7), (132049, 5, 0), (132049, 66024, 13), (9689, 1, 1),
             (756839, 377, 4), (4097, 2048, 20), (31, 3, 31), (132049, 1000, 13)]
    ins = [rand_res(rng, r) for r, _, _ in cases]
    bufs = [DeviceBuffer.from_array(a) for a in ins]
    tmps = [DeviceBuffer.words(O.words(r)) for r, _, _ in cases]
    n = len(cases)
    res = (C.c_void_p * n)(*[b.ptr for b in bufs])
    tmp = (C.c_void_p * n)(*[t.ptr for t in tmps])
    rr = np.array([c[0] for c in cases], dtype=np.uint64)
    ss = np.array([c[1] for c in cases], dtype=np.uint64)
    kk = np.array([c[2] for c in cases], dtype=np.uint64)
    P = C.POINTER(C.c_uint64)
    check(lib.gf2t_square_fold_batch(res, tmp, rr.ctypes.data_as(P), ss.ctypes.data_as(P),
                                     kk.ctypes.data_as(P), n, None))
    for (r, s, k), a, b in zip(cases, ins, bufs):
        assert np.array_equal(b.download(O.words(r)), O.square_fold_steps(a, r, s, k)), (r, s, k)


def test_batch_full_size(eng, tuned):
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    import ctypes as C
    rng = np.random.default_rng(3)
    svals = [3, 474, 1234567, 68139920]
    steps = [4, 3, 2, 1]
    nw = O.words(R136)
    ins = [rand_res(rng, R136) for _ in svals]
    bufs = [DeviceBuffer.from_array(a) for a in ins]
    tmps = [DeviceBuffer.words(nw) for _ in svals]
    n = len(svals)
    P = C.POINTER(C.c_uint64)
    rr = np.full(n, R136, dtype=np.uint64)
    ss = np.array(svals, dtype=np.uint64)
    kk = np.array(steps, dtype=np.uint64)
    check(lib.gf2t_square_fold_batch((C.c_void_p * n)(*[b.ptr for b in bufs]),
                                     (C.c_void_p * n)(*[t.ptr for t in tmps]),
                                     rr.ctypes.data_as(P), ss.ctypes.data_as(P),
                                     kk.ctypes.data_as(P), n, None))
    for s, k, a, b in zip(svals, steps, ins, bufs):
        assert np.array_equal(b.download(nw), O.square_fold_steps(a, R136, s, k)), s


def test_batch_full_size_two_squarings(eng):
    """One batched two-squaring launch at full size over residues whose s
    put the points s and 2s at very different tiles (s from 474 to
    (r - 1)/2), every residue against the oracle."""
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    import ctypes as C
    rng = np.random.default_rng(5)
    svals = [474, 65100, 1234567, 30000001, 45678901, 68139919, 68139920]
    nw = O.words(R136)
    ins = [rand_res(rng, R136) for _ in svals]
    bufs = [DeviceBuffer.from_array(a) for a in ins]
    tmps = [DeviceBuffer.words(nw) for _ in svals]
    n = len(svals)
    P = C.POINTER(C.c_uint64)
    rr = np.full(n, R136, dtype=np.uint64)
    ss = np.array(svals, dtype=np.uint64)
    kk = np.full(n, 2, dtype=np.uint64)
    check(lib.gf2t_square_fold_batch((C.c_void_p * n)(*[b.ptr for b in bufs]),
                                     (C.c_void_p * n)(*[t.ptr for t in tmps]),
                                     rr.ctypes.data_as(P), ss.ctypes.data_as(P),
                                     kk.ctypes.data_as(P), n, None))
    for s, a, b in zip(svals, ins, bufs):
        assert np.array_equal(b.download(nw), O.square_fold_steps(a, R136, s, 2)), s


def test_r_multiple_of_64(eng):
    """r a multiple of 64 takes the window loads that never read past a
    window's last word (the 12-byte form could read one word past the
    residue there): single chains of one to five steps, and a batch that
    mixes such an r with odd ones."""
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    import ctypes as C
    rng = np.random.default_rng(64)
    for r in (128, 4160, 8192, 262144):
        for s in sorted({1, 63, 64, r // 3, r // 2 - 1} & set(range(1, (r + 1) // 2))):
            a = rand_res(rng, r)
            for k in (1, 2, 5):
                assert np.array_equal(eng.square_fold(a, r, s, k), O.square_fold_steps(a, r, s, k)), (r, s, k)
    cases = [(4160, 1000, 4), (4423, 1200, 4), (262144, 70001, 2), (8192, 4095, 3)]
    ins = [rand_res(rng, r) for r, _, _ in cases]
    bufs = [DeviceBuffer.from_array(x) for x in ins]
    tmps = [DeviceBuffer.words(O.words(r)) for r, _, _ in cases]
    n = len(cases)
    P = C.POINTER(C.c_uint64)
    rr = np.array([c[0] for c in cases], dtype=np.uint64)
    ss = np.array([c[1] for c in cases], dtype=np.uint64)
    kk = np.array([c[2] for c in cases], dtype=np.uint64)
    check(lib.gf2t_square_fold_batch((C.c_void_p * n)(*[b.ptr for b in bufs]),
                                     (C.c_void_p * n)(*[t.ptr for t in tmps]),
                                     rr.ctypes.data_as(P), ss.ctypes.data_as(P),
                                     kk.ctypes.data_as(P), n, None))
    for (r, s, k), x, b in zip(cases, ins, bufs):
        assert np.array_equal(b.download(O.words(r)), O.square_fold_steps(x, r, s, k)), (r, s, k)


@pytest.mark.parametrize("r,s", [(4423, 1200), (4097, 1), (4097, 2048), (131, 65), (132049, 777)])
def test_fold_products(eng, r, s):
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    rng = np.random.default_rng(r + s)
    for L in (1, 64, r - 1, r, r + 1, r + s - 1, r + s, 2 * r - 2, 2 * r - 1):
        v = rng.integers(0, 2**64, O.words(L), dtype=np.uint64)
        if L & 63:
            v[-1] &= np.uint64((1 << (L & 63)) - 1)
        src = DeviceBuffer.from_array(v)
        out = DeviceBuffer.words(O.words(r))
        check(lib.gf2t_fold(src.ptr, L, out.ptr, r, s, None))
        assert np.array_equal(out.download(), O.reduce_bitserial(v, L, r, s)), L


@pytest.mark.parametrize("r,s", [(4423, 1200), (4097, 1), (4096 + 65, 64), (130, 64), (132049, 66024)])
def test_mulx_xor(eng, r, s):
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    rng = np.random.default_rng(s)
    for top in (0, 1):
        a = rand_res(rng, r)
        word, bit = (r - 1) >> 6, (r - 1) & 63
        a[word] = (a[word] & ~np.uint64(1 << bit)) | np.uint64(top << bit)
        d = rand_res(rng, r)
        da, dd = DeviceBuffer.from_array(a), DeviceBuffer.from_array(d)
        check(lib.gf2t_mulx_xor(dd.ptr, da.ptr, r, s, None))
        assert np.array_equal(dd.download(), d ^ O.mulx_mod(a, r, s))
