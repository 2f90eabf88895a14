"""The deep-scan workload at r = 136279841: 256 s drawn by SplitMix64
(seed 20250101) from [3, 68139920], the coarse sieve to degree 32, then
the batched deep scan of its survivors to degree 4096 with an interval
gcd every 512 degrees and per-s early exit.

- The batched squaring chain over all 256 full-size residues (4.36 GB
  resident) equals the oracle's linear spread+fold on 16 of them.
- Every certificate line: sieve lines carry their degree and mask; every
  deep-scan factor divides x^r + x^s + 1 and is irreducible (the oracle's
  check-ntl criterion); and the whole certificate's md5 is the one the
  batched and the plain one-s-at-a-time forms both produced on an MI355X
  (bench.py's c3 object, certificates_equal: true)."""
import ctypes as C
import hashlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
R = 136279841
C3_MD5 = "0e0db1573009e1ebdea27a197aad4cb1"


def test_batched_chain_256_residues_vs_oracle():
    import bench
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    device.init(0)
    svals = np.array(bench.c3_svals(20250101), dtype=np.uint64)
    n, nw, steps = len(svals), O.words(R), 64
    res = [DeviceBuffer.words(nw) for _ in range(n)]
    tmp = [DeviceBuffer.words(nw) for _ in range(n)]
    for b in res:
        check(lib.gf2t_set_x(b.ptr, nw, None))
    P64 = C.POINTER(C.c_uint64)
    rr = np.full(n, R, dtype=np.uint64)
    kk = np.full(n, steps, dtype=np.uint64)
    print(f"chain: {n} residues x {steps} squarings", flush=True)
    check(lib.gf2t_square_fold_batch((C.c_void_p * n)(*[b.ptr for b in res]),
                                     (C.c_void_p * n)(*[b.ptr for b in tmp]),
                                     rr.ctypes.data_as(P64), svals.ctypes.data_as(P64),
                                     kk.ctypes.data_as(P64), n, None))
    device.synchronize()
    pick = list(range(0, n, 16))
    got = {i: res[i].download(nw) for i in pick}
    x = O.x_residue(R)
    with ThreadPoolExecutor(16) as ex:
        want = dict(zip(pick, ex.map(lambda i: O.square_fold_steps(x, R, int(svals[i]), steps),
                                     pick)))
    for i in pick:
        assert np.array_equal(got[i], want[i]), int(svals[i])


def test_c3_certificates():
    import bench
    from mersenne_trinomial_amd import device
    device.init(0)
    svals = bench.c3_svals(20250101)
    print("c3: sieve + batched deep scan of 256 s", flush=True)
    lines, info = bench.c3_run(svals, depth=32, maxd=4096, interval=512, batch=0,
                               gcd_threads=16)
    print(f"c3: {info}", flush=True)
    assert sorted(lines) == sorted(svals)
    deep = []
    for s, ln in lines.items():
        f = ln.split()
        assert int(f[0]) == s and f[1] != "u", ln
        d = int(f[1])
        P = O.poly_from_hex(f[2][1:])
        assert O.deg(P) == d, ln
        if d > 32:
            deep.append((s, P))
        else:
            assert O.divides_trinomial_small(R, s, int(f[2][1:], 16), d), ln
    assert len(deep) == info["survivors"] > 0
    with ThreadPoolExecutor(16) as ex:
        ok = list(ex.map(lambda sp: O.verify_factor(R, sp[0], sp[1]), deep))
    assert all(ok)
    text = "".join(lines[s] + "\n" for s in sorted(lines))
    assert hashlib.md5(text.encode()).hexdigest() == C3_MD5
