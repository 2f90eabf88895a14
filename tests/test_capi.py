"""The C ABI library loads, exports every symbol include/gf2t.h declares,
and its host-only entry points behave -- no GPU needed."""
import ctypes as C

import numpy as np
import pytest

import oracle as O


def test_library_exports_every_declared_symbol():
    from mersenne_trinomial_amd import _capi
    names = _capi.declared_symbols()
    assert len(names) >= 40
    raw = C.CDLL(_capi.LIB_PATH)
    missing = [n for n in names if not hasattr(raw, n)]
    assert not missing, missing
    # and the ctypes binding covers all of them
    assert set(names) <= set(_capi._SIGS)


def test_host_only_entry_points():
    from mersenne_trinomial_amd._capi import lib
    assert lib.gf2t_version() >= 1
    for d in range(2, 21):
        assert lib.gf2t_primitive_poly(d) == O.find_primitive_poly(d)
    assert lib.gf2t_primitive_poly(1) == 0 and lib.gf2t_primitive_poly(41) == 0
    for d in (33, 34):
        q = lib.gf2t_primitive_poly(d)
        assert q >> d == 1 and O.is_irreducible_small(q, d)
    assert lib.gf2t_mul_out_words(100, 200) == O.lib().gf2o_mul_out_words(100, 200)
    b = np.zeros(64, dtype=np.uint64)
    n = C.c_int()
    assert lib.gf2t_cantor_basis(b.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(n)) == 0
    assert n.value == 64
    assert np.array_equal(b, O.cantor_basis())


def test_argument_errors_return_status_not_exit():
    from mersenne_trinomial_amd._capi import lib
    # invalid trinomials are rejected before any device work
    assert lib.gf2t_square_fold(8, 16, 31, 16, None) == -1      # 2s >= r
    assert b"2s < r" in lib.gf2t_last_error()
    assert lib.gf2t_square_fold(8, 16, 31, 0, None) == -1       # s = 0
    assert lib.gf2t_fold(8, 100, 16, 31, 3, None) == -1         # src_bits > 2r-1
    assert lib.gf2t_fft_plan_create(0, C.byref(C.c_void_p())) == -1
    assert lib.gf2t_sieve(101, 2, 37, 10, 8, None) == -1


def test_engine_basis_verify():
    from mersenne_trinomial_amd import CantorBasis
    B = CantorBasis()
    B.build()
    assert B.len == 64 and B.verify()
    B.b[5] ^= np.uint64(4)
    assert not B.verify()
