"""CPU oracle for the GF(2) trinomial smallest-factor path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import it.  The product (``mersenne_trinomial_amd``) never does.

The C restatement lives in ``gf2_oracle.c`` (built into ``liboracle.so`` by
``oracle/Makefile``); ``_ref/`` holds builds of the reference's own sources
(``sieve_ref`` and a shim over its host Cantor engine) when the reference
tree was present at build time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_u64 = C.c_uint64
_i64 = C.c_int64
_p64 = C.POINTER(C.c_uint64)


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree exists)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _load():
    if not os.path.exists(LIB_PATH):
        build(ref=False)
    lib = C.CDLL(LIB_PATH)
    sig = {
        "gf2o_clmul32": (_u64, [C.c_uint32, C.c_uint32]),
        "gf2o_fmul": (C.c_uint32, [C.c_uint32, C.c_uint32, _u64, C.c_int]),
        "gf2o_is_irreducible": (C.c_int, [_u64, C.c_int]),
        "gf2o_find_primitive_poly": (_u64, [C.c_int]),
        "gf2o_count_irreducible": (_u64, [C.c_int]),
        "gf2o_sieve": (C.c_int, [_u64, C.c_int, C.c_int, _u64, _p64, _p64, C.c_int]),
        "gf2o_brute_force": (C.c_int, [_u64, C.c_int, _u64, _p64]),
        "gf2o_divides_trinomial_small": (C.c_int, [_u64, _u64, _u64, C.c_int]),
        "gf2o_is_irreducible_small": (C.c_int, [_u64, C.c_int]),
        "gf2o_square_fold": (None, [_p64, _p64, _u64, _u64]),
        "gf2o_square_fold_steps": (None, [_p64, _u64, _u64, _u64]),
        "gf2o_reduce_fold": (None, [_p64, _u64, _u64, _u64]),
        "gf2o_reduce_bitserial": (None, [_p64, _u64, _u64, _u64]),
        "gf2o_mulx_mod": (None, [_p64, _p64, _u64, _u64]),
        "gf2o_mul_out_words": (C.c_size_t, [_u64, _u64]),
        "gf2o_mul_naive": (None, [_p64, _u64, _p64, _u64, _p64]),
        "gf2o_mul_comb": (None, [_p64, _u64, _p64, _u64, _p64]),
        "gf2o_gf64_mul_bitserial": (_u64, [_u64, _u64]),
        "gf2o_gf64_mul": (_u64, [_u64, _u64]),
        "gf2o_cantor_basis": (C.c_int, [_p64]),
        "gf2o_cantor_point": (None, [_p64, _u64, _p64]),
        "gf2o_fft_fwd": (None, [_p64, C.c_int, _p64]),
        "gf2o_fft_inv": (None, [_p64, C.c_int, _p64]),
        "gf2o_poly_eval64": (_u64, [_p64, C.c_size_t, _u64]),
        "gf2o_fft_mul": (C.c_int, [_p64, _u64, _p64, _u64, _p64]),
        "gf2o_mulmod": (C.c_int, [_p64, _u64, _p64, _u64, _u64, _u64, _p64]),
        "gf2o_residue": (_u64, [_p64, _u64, _u64, C.c_int]),
        "gf2o_find_irreducible": (_u64, [C.c_int, _u64]),
        "gf2o_edf_least": (C.c_size_t, [_p64, C.c_size_t, _u64, _p64, C.c_size_t]),
        "gf2o_scan_one": (C.c_int, [_u64, _u64, _u64, _u64, _p64, _p64, C.c_size_t, _p64]),
        "gf2o_deg": (_i64, [_p64, C.c_size_t]),
        "gf2o_gcd": (_i64, [_p64, C.c_size_t, _p64, C.c_size_t, _p64, C.c_size_t]),
        "gf2o_verify_factor": (C.c_int, [_u64, _u64, _p64, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_p64)


def words(bits: int) -> int:
    return (bits + 63) >> 6


def mul_out_words(abits: int, bbits: int) -> int:
    """Words of a product buffer (mul_out_words, src/gf2_cantor_engine.h:200-202)."""
    return lib().gf2o_mul_out_words(abits, bbits)


# ---------------------------------------------------------------- field
def fmul(a, b, q, d):
    return lib().gf2o_fmul(a, b, q, d)


def find_primitive_poly(d):
    return lib().gf2o_find_primitive_poly(d)


def count_irreducible(d):
    return lib().gf2o_count_irreducible(d)


def is_irreducible_small(mask, d):
    return bool(lib().gf2o_is_irreducible_small(mask, d))


def divides_trinomial_small(r, s, mask, d):
    return bool(lib().gf2o_divides_trinomial_small(r, s, mask, d))


# ---------------------------------------------------------------- sieve
NONE = np.uint64(0xFFFFFFFFFFFFFFFF)


def sieve(r, dmax, smax, dmin=2, nthreads=1, best=None):
    """best[0..smax] of min packed keys (d<<48)|mask, ~0 = survivor;
    returns (best, classes) with classes[d] = surviving classes."""
    if best is None:
        best = np.full(smax + 1, NONE, dtype=np.uint64)
    classes = np.zeros(max(dmax + 1, 33), dtype=np.uint64)
    rc = lib().gf2o_sieve(r, dmin, dmax, smax, _ptr(best), _ptr(classes), nthreads)
    if rc:
        raise RuntimeError(f"gf2o_sieve failed rc={rc}")
    return best, classes


def brute_force(r, depth, smax):
    best = np.full(smax + 1, NONE, dtype=np.uint64)
    lib().gf2o_brute_force(r, depth, smax, _ptr(best))
    return best


def cert_lines(best: np.ndarray):
    """'s d p<hex>' lines for every marked s (reference format,
    src/sieve_ref.cpp:163-169)."""
    out = []
    for s in np.nonzero(best != NONE)[0]:
        if s == 0:
            continue
        k = int(best[s])
        out.append(f"{s} {k >> 48} p{k & 0xFFFFFFFFFFFF:x}")
    return out


# ------------------------------------------------------------- residues
def square_fold(a: np.ndarray, r: int, s: int) -> np.ndarray:
    out = np.zeros(words(r), dtype=np.uint64)
    lib().gf2o_square_fold(_ptr(np.ascontiguousarray(a, dtype=np.uint64)), _ptr(out), r, s)
    return out


def square_fold_steps(a: np.ndarray, r: int, s: int, steps: int) -> np.ndarray:
    res = np.array(a, dtype=np.uint64, copy=True)
    lib().gf2o_square_fold_steps(_ptr(res), r, s, steps)
    return res


def reduce_fold(v: np.ndarray, bits: int, r: int, s: int) -> np.ndarray:
    buf = np.zeros(max(len(v), 2 * words(r) + 2), dtype=np.uint64)
    buf[: len(v)] = v
    lib().gf2o_reduce_fold(_ptr(buf), bits, r, s)
    return buf[: words(r)].copy()


def reduce_bitserial(v: np.ndarray, bits: int, r: int, s: int) -> np.ndarray:
    buf = np.zeros(max(len(v), 2 * words(r) + 2), dtype=np.uint64)
    buf[: len(v)] = v
    lib().gf2o_reduce_bitserial(_ptr(buf), bits, r, s)
    return buf[: words(r)].copy()


def mulx_mod(a, r, s):
    out = np.zeros(words(r), dtype=np.uint64)
    lib().gf2o_mulx_mod(_ptr(np.ascontiguousarray(a, dtype=np.uint64)), _ptr(out), r, s)
    return out


def x_residue(r: int) -> np.ndarray:
    a = np.zeros(words(r), dtype=np.uint64)
    a[0] = 2
    return a


# -------------------------------------------------------------- products
def _mul(fn, a, abits, b, bbits):
    out = np.zeros(lib().gf2o_mul_out_words(abits, bbits), dtype=np.uint64)
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    rc = fn(_ptr(a), abits, _ptr(b), bbits, _ptr(out))
    if rc:
        raise RuntimeError(f"product failed rc={rc}")
    return out


def mul_naive(a, abits, b, bbits):
    return _mul(lib().gf2o_mul_naive, a, abits, b, bbits)


def mul_comb(a, abits, b, bbits):
    return _mul(lib().gf2o_mul_comb, a, abits, b, bbits)


def fft_mul(a, abits, b, bbits):
    return _mul(lib().gf2o_fft_mul, a, abits, b, bbits)


def mulmod(a, abits, b, bbits, r, s):
    out = np.zeros(words(r), dtype=np.uint64)
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    rc = lib().gf2o_mulmod(_ptr(a), abits, _ptr(b), bbits, r, s, _ptr(out))
    if rc:
        raise RuntimeError(f"mulmod failed rc={rc}")
    return out


# ---------------------------------------------------------- Cantor / FFT
def gf64_mul(a, b):
    return lib().gf2o_gf64_mul(a, b)


def gf64_mul_bitserial(a, b):
    return lib().gf2o_gf64_mul_bitserial(a, b)


def cantor_basis() -> np.ndarray:
    b = np.zeros(64, dtype=np.uint64)
    n = lib().gf2o_cantor_basis(_ptr(b))
    if n != 64:
        raise RuntimeError(f"Cantor chain length {n}")
    return b


def cantor_point(basis, idx):
    out = np.zeros(1, dtype=np.uint64)
    lib().gf2o_cantor_point(_ptr(basis), idx, _ptr(out))
    return int(out[0])


def fft_fwd(f: np.ndarray, m: int) -> np.ndarray:
    f = np.array(f, dtype=np.uint64, copy=True)
    assert len(f) == 1 << m
    lib().gf2o_fft_fwd(_ptr(f), m, _ptr(cantor_basis()))
    return f


def fft_inv(f: np.ndarray, m: int) -> np.ndarray:
    f = np.array(f, dtype=np.uint64, copy=True)
    assert len(f) == 1 << m
    lib().gf2o_fft_inv(_ptr(f), m, _ptr(cantor_basis()))
    return f


def poly_eval64(f: np.ndarray, pt: int) -> int:
    f = np.ascontiguousarray(f, dtype=np.uint64)
    return lib().gf2o_poly_eval64(_ptr(f), len(f), pt)


def residue(a: np.ndarray, abits: int, q: int, d: int) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return lib().gf2o_residue(_ptr(a), abits, q, d)


def find_irreducible(d: int, seed: int) -> int:
    return lib().gf2o_find_irreducible(d, seed)


# --------------------------------------------------------------- gcd etc
def deg(w: np.ndarray) -> int:
    w = np.ascontiguousarray(w, dtype=np.uint64)
    return lib().gf2o_deg(_ptr(w), len(w))


def gcd(a: np.ndarray, b: np.ndarray):
    """(degree, gcd words) -- degree -1 iff both are zero."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    cap = max(len(a), len(b)) + 1
    g = np.zeros(cap, dtype=np.uint64)
    d = lib().gf2o_gcd(_ptr(a), len(a), _ptr(b), len(b), _ptr(g), cap)
    return d, g[: max(1, words(d + 1))] if d >= 0 else g[:1]


def verify_factor(r, s, P: np.ndarray) -> bool:
    """1 iff P divides x^r+x^s+1 and (for deg <= 20000) is irreducible --
    the check-ntl verifier's criterion (src/check-ntl.c:17-58)."""
    P = np.ascontiguousarray(P, dtype=np.uint64)
    return bool(lib().gf2o_verify_factor(r, s, _ptr(P), len(P)))


def edf_least(t: np.ndarray, d: int) -> np.ndarray:
    t = np.ascontiguousarray(t, dtype=np.uint64)
    cap = words(d + 1)
    out = np.zeros(cap, dtype=np.uint64)
    n = lib().gf2o_edf_least(_ptr(t), len(t), d, _ptr(out), cap)
    if n == 0:
        raise RuntimeError("no factor found")
    return out


def scan_one(r: int, s: int, k0: int, maxd: int):
    """(kind, d, mask_hex, gcd_degree): kind 0 = not a survivor,
    1 = found, 2 = no factor up to maxd."""
    cap = words(maxd + 1) + 1
    mask = np.zeros(cap, dtype=np.uint64)
    d = np.zeros(1, dtype=np.uint64)
    gd = np.zeros(1, dtype=np.uint64)
    kind = lib().gf2o_scan_one(r, s, k0, maxd, _ptr(d), _ptr(mask), cap, _ptr(gd))
    return kind, int(d[0]), (poly_hex(mask) if kind == 1 else None), int(gd[0])


def trinomial(r: int, s: int) -> np.ndarray:
    t = np.zeros(words(r + 1), dtype=np.uint64)
    for e in (0, s, r):
        t[e >> 6] |= np.uint64(1 << (e & 63))
    return t


def poly_from_hex(h: str) -> np.ndarray:
    v = int(h, 16)
    n = max(1, (v.bit_length() + 63) // 64)
    return np.array([(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(n)], dtype=np.uint64)


def poly_hex(w: np.ndarray) -> str:
    v = 0
    for i, x in enumerate(np.asarray(w, dtype=np.uint64)):
        v |= int(x) << (64 * i)
    return format(v, "x")


# ------------------------------------------------------- reference (_ref)
def have_ref() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libgf2ref.so"))


_ref = None


def ref():
    global _ref
    if _ref is None:
        lib_ = C.CDLL(os.path.join(REF_DIR, "libgf2ref.so"))
        for name, res, args in [
            ("ref_cantor_basis", C.c_int, [_p64]),
            ("ref_gf64_mul", _u64, [_u64, _u64]),
            ("ref_fft_fwd", None, [_p64, C.c_int]),
            ("ref_fft_inv", None, [_p64, C.c_int]),
            ("ref_fft_mul", None, [_p64, _u64, _p64, _u64, _p64]),
            ("ref_gcd_naive", _i64, [_p64, C.c_size_t, _p64, C.c_size_t, _p64, C.c_size_t]),
            ("ref_poly_hex", C.c_int, [_p64, C.c_size_t, C.c_char_p, C.c_size_t]),
        ]:
            fn = getattr(lib_, name)
            fn.restype = res
            fn.argtypes = args
        _ref = lib_
    return _ref


def ref_sieve_lines(r, depth, smax):
    """Run the reference's sieve_ref (built in _ref/) and return its lines."""
    exe = os.path.join(REF_DIR, "sieve_ref")
    out = subprocess.run([exe, "run", str(r), str(depth), str(smax)],
                         check=True, capture_output=True, text=True).stdout
    return out.splitlines()
