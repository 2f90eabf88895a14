"""Python mirror of the reference engine API over the C ABI.

Names, argument meaning and error behaviour follow the reference's
engine structs so code written against them reads the same:

  CantorBasis  src/gf2_cantor_engine.h:74-121
  DevFFT       CudaFFT, src/gf2_cantor_cuda.h:361-521
  TrinGPU      src/gf2_trinomial_cuda.h:87-178
  run_gpu      the coarse sieve driver, src/coarse_sieve.cu:1119-1225

The reference aborts with exit(1) on bad input; here the C ABI returns a
status and these wrappers raise GF2TError / ValueError instead.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import GF2TError, SieveStats, check, lib
from .device import DeviceBuffer, init

U64_MAX = 0xFFFFFFFFFFFFFFFF


def bits_to_words(bits: int) -> int:
    """src/gf2_cantor_engine.h:198"""
    return (int(bits) + 63) >> 6


def mul_out_words(abits: int, bbits: int) -> int:
    """src/gf2_cantor_engine.h:201-203"""
    return int(lib.gf2t_mul_out_words(abits, bbits))


def _clmul(a: int, b: int) -> int:
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        b >>= 1
    return r


def _gf64_mul(a: int, b: int) -> int:
    p = _clmul(a, b)
    for bit in range(126, 63, -1):
        if (p >> bit) & 1:
            p ^= 0x1B << (bit - 64) | (1 << bit)
    return p


class CantorBasis:
    """b[0] = 1, b[i]^2 + b[i] = b[i-1] in GF(2^64) (len = 64)."""

    def __init__(self):
        self.b = np.zeros(64, dtype=np.uint64)
        self.len = 0

    def build(self) -> None:
        n = C.c_int(0)
        check(lib.gf2t_cantor_basis(self.b.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(n)),
              "gf2t_cantor_basis")
        self.len = n.value

    def verify(self) -> bool:
        b = [int(x) for x in self.b[: self.len]]
        if not b or b[0] != 1:
            return False
        for i in range(1, len(b)):
            if _gf64_mul(b[i], b[i]) ^ b[i] != b[i - 1]:
                return False
        # linear independence over GF(2)
        piv = {}
        for v in b:
            while v:
                top = v.bit_length() - 1
                if top in piv:
                    v ^= piv[top]
                else:
                    piv[top] = v
                    break
            if not v:
                return False
        return True


class DevFFT:
    """Device Cantor transform of 2^m GF(2^64) elements (CudaFFT)."""

    def __init__(self):
        self.m = 0
        self.n = 0
        self._plan = None

    def init(self, m: int, basis: CantorBasis | None = None) -> None:
        init()
        self.fini()
        p = C.c_void_p()
        check(lib.gf2t_fft_plan_create(int(m), C.byref(p)), "gf2t_fft_plan_create")
        self._plan = p.value
        self.m = int(m)
        self.n = 1 << self.m

    def passes(self):
        f, i = C.c_int(), C.c_int()
        check(lib.gf2t_fft_plan_passes(self._plan, C.byref(f), C.byref(i)))
        return f.value, i.value

    def fwd(self, df, stream=None) -> None:
        check(lib.gf2t_fft_fwd(self._plan, _ptr(df), stream), "gf2t_fft_fwd")

    def inv(self, df, stream=None) -> None:
        check(lib.gf2t_fft_inv(self._plan, _ptr(df), stream), "gf2t_fft_inv")

    def fini(self) -> None:
        if self._plan:
            lib.gf2t_fft_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.fini()
        except Exception:
            pass


def _ptr(x) -> int:
    return x.ptr if isinstance(x, DeviceBuffer) else int(x)


class TrinGPU:
    """Arithmetic in GF(2)[x]/(x^r + x^s + 1) on the device.

    Owns dA (F_k), dACC (interval product), dT0 (2r-bit product) and the
    product workspace, like the reference context."""

    def __init__(self):
        self.r = 0
        self.s = 0

    def init(self, r: int) -> None:
        init()
        self.r = int(r)
        self.s = 0
        self.nw = bits_to_words(self.r)
        self.nw2 = 2 * self.nw
        self.dA = DeviceBuffer.words(self.nw)
        self.dACC = DeviceBuffer.words(self.nw)
        self.dT0 = DeviceBuffer.words(self.nw2 + 2)
        self.dT1 = DeviceBuffer.words(self.nw)
        ws = int(lib.gf2t_mul_workspace_bytes(self.r, self.r))
        self.dW = DeviceBuffer(ws)
        ca = (self.r + 31) >> 5
        m = 1
        while (1 << m) < 2 * ca - 1:
            m += 1
        self.m = m

    def set_s(self, s: int) -> None:
        """requires 1 <= s <= (r-1)/2 (src/gf2_trinomial_cuda.h:118-123)"""
        if not (s >= 1 and 2 * s < self.r):
            raise ValueError("need 1 <= s <= (r-1)/2")
        self.s = int(s)

    def set_x(self) -> None:
        check(lib.gf2t_set_x(self.dA.ptr, self.nw, None), "gf2t_set_x")

    def reset_acc(self) -> None:
        check(lib.gf2t_set_one(self.dACC.ptr, self.nw, None), "gf2t_set_one")

    def reduce(self, dst, L: int) -> None:
        """reduce dT0 (bit length L) mod T into dst (nw words)"""
        check(lib.gf2t_fold(self.dT0.ptr, int(L), _ptr(dst), self.r, self.s, None), "gf2t_fold")

    def square_A(self) -> None:
        check(lib.gf2t_square_fold(self.dA.ptr, self.dT1.ptr, self.r, self.s, None),
              "gf2t_square_fold")
        self.dA, self.dT1 = self.dT1, self.dA

    def square_A_steps(self, k: int) -> None:
        check(lib.gf2t_square_fold_steps(self.dA.ptr, self.dT1.ptr, self.r, self.s, int(k), None),
              "gf2t_square_fold_steps")

    def mul_dev(self, dx, dy) -> None:
        """dT0 <- x * y (full 2r-1-bit product)"""
        check(lib.gf2t_mul(_ptr(dx), self.r, _ptr(dy), self.r, self.dT0.ptr, self.dW.ptr, None),
              "gf2t_mul")

    def accumulate(self) -> None:
        """dACC <- dACC * (dA + x) mod T"""
        check(lib.gf2t_xor_x(self.dA.ptr, None))
        check(lib.gf2t_mulmod(self.dACC.ptr, self.r, self.dA.ptr, self.r, self.r, self.s,
                              self.dACC.ptr, self.dW.ptr, None), "gf2t_mulmod")
        check(lib.gf2t_xor_x(self.dA.ptr, None))

    def A(self) -> np.ndarray:
        return self.dA.download(self.nw)

    def ACC(self) -> np.ndarray:
        return self.dACC.download(self.nw)

    def fini(self) -> None:
        for name in ("dA", "dACC", "dT0", "dT1", "dW"):
            b = getattr(self, name, None)
            if b is not None:
                b.free()


# ----------------------------------------------------------------- helpers
def square_fold(a: np.ndarray, r: int, s: int, steps: int = 1) -> np.ndarray:
    """a^(2^steps) mod x^r + x^s + 1 on the device (host in/out)."""
    nw = bits_to_words(r)
    d = DeviceBuffer.from_array(np.asarray(a, dtype=np.uint64)[:nw])
    t = DeviceBuffer.words(nw)
    check(lib.gf2t_square_fold_steps(d.ptr, t.ptr, r, s, int(steps), None),
          "gf2t_square_fold_steps")
    return d.download(nw)


def mul(a: np.ndarray, abits: int, b: np.ndarray, bbits: int) -> np.ndarray:
    """Full GF(2)[x] product on the device (gf2x_fft_mul)."""
    init()
    da = DeviceBuffer.from_array(np.asarray(a, dtype=np.uint64)[: bits_to_words(abits)])
    db = DeviceBuffer.from_array(np.asarray(b, dtype=np.uint64)[: bits_to_words(bbits)])
    ow = mul_out_words(abits, bbits)
    dout = DeviceBuffer.words(ow)
    dw = DeviceBuffer(max(8, int(lib.gf2t_mul_workspace_bytes(abits, bbits))))
    check(lib.gf2t_mul(da.ptr, abits, db.ptr, bbits, dout.ptr, dw.ptr, None), "gf2t_mul")
    return dout.download(ow)


def mulmod(a, abits, b, bbits, r, s) -> np.ndarray:
    init()
    nw = bits_to_words(r)
    da = DeviceBuffer.from_array(np.asarray(a, dtype=np.uint64)[: bits_to_words(abits)])
    db = DeviceBuffer.from_array(np.asarray(b, dtype=np.uint64)[: bits_to_words(bbits)])
    dout = DeviceBuffer.words(nw)
    dw = DeviceBuffer(max(8, int(lib.gf2t_mul_workspace_bytes(abits, bbits))))
    check(lib.gf2t_mulmod(da.ptr, abits, db.ptr, bbits, r, s, dout.ptr, dw.ptr, None),
          "gf2t_mulmod")
    return dout.download(nw)


def fft_fwd(f: np.ndarray, m: int, inverse: bool = False) -> np.ndarray:
    init()
    F = DevFFT()
    F.init(m)
    d = DeviceBuffer.from_array(np.asarray(f, dtype=np.uint64))
    (F.inv if inverse else F.fwd)(d)
    out = d.download(1 << m)
    F.fini()
    return out


def primitive_poly(d: int) -> int:
    return int(lib.gf2t_primitive_poly(d))


def run_gpu(r: int, depth: int, smax: int, best: np.ndarray | None = None,
            min_depth: int = 2):
    """Coarse sieve of all s in [1, smax] over degrees min_depth..depth.

    Returns (best, stats): best[s] = (d<<48)|mask of the least factor of
    degree <= depth (U64_MAX = survivor), stats per degree.  A `best`
    image from an earlier run seeds the minimum (the reference's --load,
    src/coarse_sieve.cu:1128-1140)."""
    init()
    if best is None:
        best = np.full(smax + 1, U64_MAX, dtype=np.uint64)
    else:
        best = np.ascontiguousarray(best, dtype=np.uint64)
        if best.size != smax + 1:
            raise ValueError("best must have smax+1 entries")
    dbest = DeviceBuffer.from_array(best)
    stats = (SieveStats * 37)()
    rc = lib.gf2t_sieve(r, int(min_depth), int(depth), int(smax), dbest.ptr, stats)
    check(rc, "gf2t_sieve")
    out = dbest.download(smax + 1)
    dbest.free()
    st = [stats[d] for d in range(min_depth, depth + 1)]
    return out, st


__all__ = ["CantorBasis", "DevFFT", "TrinGPU", "GF2TError", "run_gpu", "mul", "mulmod",
           "square_fold", "fft_fwd", "bits_to_words", "mul_out_words", "primitive_poly",
           "U64_MAX"]
