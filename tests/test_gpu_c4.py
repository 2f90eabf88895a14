"""The Cantor-product batch: 64 seeded operand pairs of 2^16..2^27 bits
(log-uniform, i.i.d. bits), each product folded mod x^136279841 + x^s + 1
for a seeded s -- the workload bench.py --mode fft times.  Products with
up to 2^22 bits in total are compared with the oracle's product; larger
ones through residues modulo two degree-61 irreducibles (the reference's
bigtest check, src/cantor_ref.cpp:159-224).  Every fold is compared with
the oracle's linear fold of the verified product."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
R136 = 136279841


def _clmul_mod(x: int, y: int, q: int, d: int) -> int:
    p = 0
    while y:
        if y & 1:
            p ^= x
        x <<= 1
        y >>= 1
    for b in range(p.bit_length() - 1, d - 1, -1):
        if (p >> b) & 1:
            p ^= q << (b - d)
    return p


def test_c4_batch_products_and_folds():
    import bench
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    device.init(0)
    qs = [O.find_irreducible(61, seed) for seed in (0x1234567, 0xC0FFEE)]
    pairs = bench.c4_pairs(20250101)
    assert len(pairs) == 64
    sizes = sorted(ab + bb for _, ab, _, bb, _ in pairs)
    assert sizes[0] >= 2 * (1 << 16) - 2 and sizes[-1] <= 2 * (1 << 27)
    direct = 0
    folds = []
    for i, (a, ab, b, bb, s) in enumerate(pairs):
        print(f"pair {i}: {ab} x {bb} bits, s={s}", flush=True)
        p = M.mul(a, ab, b, bb)
        if ab + bb <= 1 << 22:
            assert np.array_equal(p, O.fft_mul(a, ab, b, bb)), i
            direct += 1
        else:
            for q in qs:
                want = _clmul_mod(O.residue(a, ab, q, 61), O.residue(b, bb, q, 61), q, 61)
                assert O.residue(p, ab + bb, q, 61) == want, (i, hex(q))
        got = M.mulmod(a, ab, b, bb, R136, s)
        assert np.array_equal(got, O.reduce_fold(p, ab + bb - 1, R136, s)), i
        folds.append(got)
    assert direct >= 10
    # the same 64 products as one variable-size batch (gf2t_mulmod_batch:
    # grouped launches over every pair's passes)
    run_batch(pairs, folds)


def run_batch(pairs, want):
    import ctypes as C
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    n = len(pairs)
    P64 = C.POINTER(C.c_uint64)
    da = [DeviceBuffer.from_array(a) for a, _, _, _, _ in pairs]
    db = [DeviceBuffer.from_array(b) for _, _, b, _, _ in pairs]
    nw = O.words(R136)
    out = [DeviceBuffer.words(nw) for _ in range(n)]
    for o in out:
        o.fill(0x77)
    ab = np.array([p[1] for p in pairs], dtype=np.uint64)
    bb = np.array([p[3] for p in pairs], dtype=np.uint64)
    ss = np.array([p[4] for p in pairs], dtype=np.uint64)
    wb = lib.gf2t_mulmod_batch_workspace_bytes(n, ab.ctypes.data_as(P64), bb.ctypes.data_as(P64))
    work = DeviceBuffer(wb)
    vp = C.c_void_p * n
    print(f"batch of {n}: workspace {wb / 2**30:.2f} GiB", flush=True)
    for rep in range(2):          # the second call reuses the workspace
        check(lib.gf2t_mulmod_batch(n, vp(*[x.ptr for x in da]), ab.ctypes.data_as(P64),
                                    vp(*[x.ptr for x in db]), bb.ctypes.data_as(P64), R136,
                                    ss.ctypes.data_as(P64), vp(*[x.ptr for x in out]), work.ptr,
                                    None), "gf2t_mulmod_batch")
        device.synchronize()
        for i in range(n):
            assert np.array_equal(out[i].download(nw), want[i]), (rep, i)


def test_batch_edge_pairs():
    """Zero operands, one-chunk operands, and a mix of tiny and large pairs
    in one batch equal one-at-a-time products."""
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    device.init(0)
    rng = np.random.default_rng(7)
    pairs = []
    for ab, bb in [(1, 1), (0, 500), (31, 5), (64, 64), (3000, 100000), (1 << 20, 1 << 18),
                   (200, 1), (77, 0)]:
        a = rng.integers(0, 2**64, (ab + 63) // 64, dtype=np.uint64)
        b = rng.integers(0, 2**64, (bb + 63) // 64, dtype=np.uint64)
        if ab & 63:
            a[-1] &= np.uint64((1 << (ab & 63)) - 1)
        if bb & 63:
            b[-1] &= np.uint64((1 << (bb & 63)) - 1)
        pairs.append((a, ab, b, bb, int(rng.integers(1, R136 // 2))))
    want = [M.mulmod(a, ab, b, bb, R136, s) for a, ab, b, bb, s in pairs]
    run_batch(pairs, want)
