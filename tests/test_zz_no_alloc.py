"""gf2t_mulmod_noalloc, the modmul whose workspace is the library's own
device array: bit-identical to gf2t_mulmod on two calls in a row (the
largest product the array holds, then a smaller one), the size limit
reported as an error, and a zero operand."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GF2T_ERANGE = -4


def _poly(rng, bits):
    a = rng.integers(0, 2**64, (bits + 63) // 64, dtype=np.uint64)
    if bits & 63:
        a[-1] &= np.uint64((1 << (bits & 63)) - 1)
    return a


def _both(lib, device, r, s, abits, bbits, seed):
    rng = np.random.default_rng(seed)
    a, b = _poly(rng, abits), _poly(rng, bbits)
    da = device.DeviceBuffer.from_array(a) if abits else device.DeviceBuffer.words(1)
    db = device.DeviceBuffer.from_array(b) if bbits else device.DeviceBuffer.words(1)
    nw = O.words(r)
    ref = device.DeviceBuffer.words(nw)
    got = device.DeviceBuffer.words(nw)
    got.fill(0xA5)
    work = device.DeviceBuffer(max(8, lib.gf2t_mul_workspace_bytes(abits, bbits)))
    assert lib.gf2t_mulmod(da.ptr, abits, db.ptr, bbits, r, s, ref.ptr, work.ptr, None) == 0
    assert lib.gf2t_mulmod_noalloc(da.ptr, abits, db.ptr, bbits, r, s, got.ptr, None) == 0
    device.synchronize()
    return a, b, ref.download(nw), got.download(nw)


def test_noalloc_equals_mulmod_two_calls():
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import lib
    device.init(0)
    # first the largest product the scratch holds: 2^21 + 2^21 chunks
    # -> 2^22 transform points; then a smaller one in the same array
    for r, s, ab, bb, seed in [((1 << 26) + 1, 12345, 1 << 26, 1 << 26, 1),
                               (132049, 1201, 132049, 132049, 2)]:
        print(f"noalloc mulmod r={r} s={s} {ab}x{bb} bits", flush=True)
        a, b, ref, got = _both(lib, device, r, s, ab, bb, seed)
        assert np.array_equal(got, ref), (r, s)
        if r < 200000:
            assert np.array_equal(got, O.mulmod(a, ab, b, bb, r, s))


def test_noalloc_limits_and_zero_operand():
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import lib
    device.init(0)
    r = (1 << 27) + 1
    big = device.DeviceBuffer.words((1 << 20) + 2)
    out = device.DeviceBuffer.words(O.words(r))
    # at the limit: 2^21 + 1 and 2^21 chunks need exactly 2^22 points ...
    ok = lib.gf2t_mulmod_noalloc(big.ptr, (1 << 26) + 32, big.ptr, 1 << 26, r, 3, out.ptr, None)
    assert ok == 0
    # ... one chunk more does not fit
    rc = lib.gf2t_mulmod_noalloc(big.ptr, (1 << 26) + 64, big.ptr, 1 << 26, r, 3, out.ptr, None)
    assert rc == GF2T_ERANGE
    assert b"2^22" in lib.gf2t_last_error()
    # a product beyond 2r - 1 bits, and s out of range
    assert lib.gf2t_mulmod_noalloc(big.ptr, 200, big.ptr, 200, 101, 3, out.ptr, None) != 0
    assert lib.gf2t_mulmod_noalloc(big.ptr, 64, big.ptr, 64, 101, 51, out.ptr, None) != 0
    # a zero operand gives the zero residue, like gf2t_mulmod
    _, _, ref, got = _both(lib, device, 4423, 505, 0, 4000, 3)
    assert np.array_equal(got, ref) and not got.any()
    # and the array still serves a full product afterwards
    _, _, ref, got = _both(lib, device, 4423, 505, 4423, 4423, 4)
    assert np.array_equal(got, ref)
