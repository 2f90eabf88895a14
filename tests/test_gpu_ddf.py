"""Blocked distinct-degree factorization on slabs (h-table, Horner
product) against the oracle's squaring and products, and the deep scan
in its plain (one product per degree) and blocked forms."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import lib
    device.init(0)
    return device, lib


def frob(r, s, k):
    return O.square_fold_steps(O.x_residue(r), r, s, k)


def esym(ys, r, s):
    """e_1..e_m of the residues ys modulo x^r + x^s + 1 (oracle products)."""
    nw = O.words(r)
    e = [np.zeros(nw, np.uint64) for _ in range(len(ys) + 1)]
    e[0][0] = 1
    for t, y in enumerate(ys):
        for j in range(t + 1, 0, -1):
            e[j] = e[j] ^ O.mulmod(e[j - 1], r, y, r, r, s)
    return e[1:]


def upload(device, words_list, pitch):
    n = len(words_list)
    host = np.zeros(n * pitch, np.uint64)
    for i, w in enumerate(words_list):
        host[i * pitch:i * pitch + len(w)] = w
    return device.DeviceBuffer.from_array(host)


@pytest.mark.parametrize("r,m,kstart", [(4423, 1, 0), (4423, 4, 5), (4423, 7, 33),
                                        (132049, 16, 21), (9689, 63, 2), (9689, 64, 3),
                                        (300007, 4, 9)])
def test_init_and_horner_vs_oracle(dev, r, m, kstart):
    device, lib = dev
    svals = [1, 2, 63, 64, 65, (r - 1) // 2, r // 3][: 5 if r > 100000 else 7]
    n = len(svals)
    nw = O.words(r)
    pitch = (nw + 31) & ~31
    H = device.DeviceBuffer.words(n * m * pitch)
    Halt = device.DeviceBuffer.words(n * m * pitch)
    Cb = device.DeviceBuffer.words(n * pitch)
    S = device.DeviceBuffer.from_array(np.array(svals, np.uint64))
    assert lib.gf2t_ddf_init_slab(H.ptr, Halt.ptr, pitch, m, r, S.ptr, kstart, n, None) == 0
    assert lib.gf2t_horner_slab(H.ptr, pitch, m, r, S.ptr, Cb.ptr, pitch, n, None) == 0
    h = H.download().reshape(n, m, pitch)
    c = Cb.download().reshape(n, pitch)
    for i, s in enumerate(svals):
        ys = [frob(r, s, kstart + t) for t in range(m)]
        want = esym(ys, r, s)
        for j in range(m):
            assert np.array_equal(h[i, j, :nw], want[j]), (s, j)
        prod = np.zeros(nw, np.uint64)
        prod[0] = 1
        for y in ys:
            t = y.copy()
            t[0] ^= np.uint64(2)
            prod = O.mulmod(prod, r, t, r, r, s)
        assert np.array_equal(c[i, :nw], prod), s
    # one Frobenius step of the whole table = the table of block kstart+1
    assert lib.gf2t_square_fold_slab_g(H.ptr, Halt.ptr, pitch, r, S.ptr, n * m, m, None) == 0
    h2 = Halt.download().reshape(n, m, pitch)
    for i, s in enumerate(svals):
        for j in range(m):
            assert np.array_equal(h2[i, j, :nw], O.square_fold(h[i, j, :nw], r, s))


@pytest.mark.parametrize("r,group", [(4423, 1), (300007, 2), (756839, 3)])
def test_two_squarings_on_a_slab(dev, r, group):
    """gf2t_square2_fold_slab_g = two squarings, residue y with s[y / group]"""
    device, lib = dev
    svals = [1, 3, min(4000, (r - 1) // 4), (r - 1) // 2 - 5, (r - 1) // 2]   # 2s < r
    n = len(svals) * group
    nw = O.words(r)
    pitch = (nw + 31) & ~31
    rng = np.random.default_rng(r)
    a = [rng.integers(0, 2**64, nw, dtype=np.uint64) for _ in range(n)]
    for v in a:
        if r & 63:
            v[-1] &= np.uint64((1 << (r & 63)) - 1)
    A = upload(device, a, pitch)
    B = device.DeviceBuffer.words(n * pitch)
    S = device.DeviceBuffer.from_array(np.array(svals, np.uint64))
    assert lib.gf2t_square2_fold_slab_g(A.ptr, B.ptr, pitch, r, S.ptr, n, group, None) == 0
    got = B.download().reshape(n, pitch)
    for y in range(n):
        s = svals[y // group]
        assert np.array_equal(got[y, :nw], O.square_fold_steps(a[y], r, s, 2)), (y, s)
    # odd pitch: rejected (the kernel stores 16-byte pairs)
    assert lib.gf2t_square2_fold_slab_g(A.ptr, B.ptr, pitch + 1, r, S.ptr, 1, 1, None) != 0
    if r < 5:
        return
    # the work-unit form with the groups' layouts in a plan: the same
    # result; then the plan reset to the groups in reverse order (as after
    # a compaction) and a slab of fewer groups
    import ctypes as C
    from mersenne_trinomial_amd._capi import _p64
    hs = np.array(svals, np.uint64)
    plan = C.c_void_p()
    assert lib.gf2t_sq2_slab_plan_create(r, hs.ctypes.data_as(_p64), len(svals),
                                         C.byref(plan)) == 0
    try:
        Bp = device.DeviceBuffer.words(n * pitch)
        assert lib.gf2t_square2_fold_slab_p(plan, A.ptr, Bp.ptr, pitch, n, group, None) == 0
        assert np.array_equal(Bp.download().reshape(n, pitch)[:, :nw], got[:, :nw])
        rev = hs[::-1].copy()
        assert lib.gf2t_sq2_slab_plan_set(plan, rev.ctypes.data_as(_p64), len(rev), None) == 0
        n2 = n - group
        assert lib.gf2t_square2_fold_slab_p(plan, A.ptr, Bp.ptr, pitch, n2, group, None) == 0
        got2 = Bp.download().reshape(n, pitch)
        for y in range(n2):
            s = int(rev[y // group])
            assert np.array_equal(got2[y, :nw], O.square_fold_steps(a[y], r, s, 2)), (y, s)
        # more residues than the plan's groups cover: rejected
        assert lib.gf2t_square2_fold_slab_p(plan, A.ptr, Bp.ptr, pitch, n + 1, group, None) != 0
        device.synchronize()
    finally:
        assert lib.gf2t_sq2_slab_plan_destroy(plan) == 0


def test_slab_helpers(dev):
    device, lib = dev
    r, n = 4423, 5
    nw = O.words(r)
    pitch = nw + 3
    rng = np.random.default_rng(4)
    svals = np.array([1, 7, 64, 1000, 2211], np.uint64)
    a = [rng.integers(0, 2**64, nw, dtype=np.uint64) for _ in range(n)]
    b = [rng.integers(0, 2**64, nw, dtype=np.uint64) for _ in range(n)]
    for v in a + b:
        v[-1] &= np.uint64((1 << (r & 63)) - 1)
    A, Bb = upload(device, a, pitch), upload(device, b, pitch)
    S = device.DeviceBuffer.from_array(svals)
    assert lib.gf2t_mulx_xor_slab(A.ptr, pitch, Bb.ptr, pitch, r, S.ptr, n, None) == 0
    got = A.download().reshape(n, pitch)
    for i in range(n):
        want = a[i] ^ O.mulx_mod(b[i], r, int(svals[i]))
        assert np.array_equal(got[i, :nw], want)
    assert lib.gf2t_set_one_slab(A.ptr, pitch, nw, n, None) == 0
    one = A.download().reshape(n, pitch)
    assert np.all(one[:, 0] == 1) and not np.any(one[:, 1:nw])


@pytest.mark.parametrize("block_m", [1, 4, 16])
def test_scan_blocked_equals_plain_equals_oracle(block_m):
    import mersenne_trinomial_amd as M
    r, k0, maxd = 4423, 5, 200
    want = {}
    for s in range(2, 351):
        kind, d, hx, _ = O.scan_one(r, s, k0, maxd)
        if kind:
            want[s] = f"{s} {d} p{hx}" if kind == 1 else f"{s} u"
    got = M.scan(r, sorted(want), k0=k0, maxd=maxd, L0=32, Lmax=256, block_m=block_m)
    assert {o.s: o.line() for o in got} == want


def test_scan_blocked_r132049_survivors():
    import mersenne_trinomial_amd as M
    r = 132049
    best, _ = O.sieve(r, 20, 66024)
    surv = [int(s) for s in np.nonzero(best == O.NONE)[0] if s >= 1][:64]
    a = {o.s: o.line() for o in M.scan(r, surv, k0=20, maxd=400, block_m=16)}
    b = {o.s: o.line() for o in M.scan(r, surv, k0=20, maxd=400, block_m=1)}
    assert a == b
    for s, line in a.items():
        parts = line.split()
        if len(parts) == 3:
            assert O.verify_factor(r, s, O.poly_from_hex(parts[2][1:]))
