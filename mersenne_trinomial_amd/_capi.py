"""ctypes binding of the C ABI declared in include/gf2t.h (libgf2t.so).

The library is the only compute path: if it is missing or fails to load,
importing this module raises, so nothing can silently run elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgf2t.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "gf2t.h")

_u64 = C.c_uint64
_p64 = C.POINTER(C.c_uint64)
_vp = C.c_void_p
_sz = C.c_size_t


class GF2TError(RuntimeError):
    """A C-ABI call returned a nonzero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"gf2t error {code}: {msg}")
        self.code = code


class ScanParams(C.Structure):
    _fields_ = [("k0", _u64), ("maxd", _u64), ("L0", _u64), ("Lmax", _u64),
                ("gcd_threads", C.c_int), ("batch", C.c_int), ("verbose", C.c_int),
                ("block_m", C.c_int), ("past_bound", C.c_int),
                ("progress", C.c_void_p), ("progress_user", C.c_void_p)]


class ScanResult(C.Structure):
    _fields_ = [("s", _u64), ("kind", C.c_int), ("d", _u64), ("rlo", _u64), ("rhi", _u64),
                ("gcd_degree", _u64), ("hex_len", _sz)]


class SieveStats(C.Structure):
    _fields_ = [("d", C.c_int), ("q", _u64), ("classes", _u64), ("marking", _u64),
                ("heavy", _u64), ("ms_table", C.c_float), ("ms_classes", C.c_float),
                ("ms_heavy", C.c_float)]


_SIGS = {
    "gf2t_last_error": (C.c_char_p, []),
    "gf2t_version": (C.c_int, []),
    "gf2t_init": (C.c_int, [C.c_int]),
    "gf2t_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "gf2t_device_name": (C.c_int, [C.c_char_p, _sz]),
    "gf2t_malloc": (C.c_int, [C.POINTER(_vp), _sz]),
    "gf2t_free": (C.c_int, [_vp]),
    "gf2t_mem_info": (C.c_int, [C.POINTER(_sz), C.POINTER(_sz)]),
    "gf2t_memcpy_h2d": (C.c_int, [_vp, _vp, _sz, _vp]),
    "gf2t_memcpy_d2h": (C.c_int, [_vp, _vp, _sz, _vp]),
    "gf2t_memcpy_d2d": (C.c_int, [_vp, _vp, _sz, _vp]),
    "gf2t_memset": (C.c_int, [_vp, C.c_int, _sz, _vp]),
    "gf2t_stream_create": (C.c_int, [C.POINTER(_vp)]),
    "gf2t_stream_destroy": (C.c_int, [_vp]),
    "gf2t_stream_sync": (C.c_int, [_vp]),
    "gf2t_device_sync": (C.c_int, []),
    "gf2t_event_create": (C.c_int, [C.POINTER(_vp)]),
    "gf2t_event_destroy": (C.c_int, [_vp]),
    "gf2t_event_record": (C.c_int, [_vp, _vp]),
    "gf2t_event_elapsed_ms": (C.c_int, [_vp, _vp, C.POINTER(C.c_float)]),
    "gf2t_stream_copy": (C.c_int, [_vp, _vp, _sz, _vp]),
    "gf2t_graph_begin": (C.c_int, [_vp]),
    "gf2t_graph_end": (C.c_int, [_vp, C.POINTER(_vp)]),
    "gf2t_graph_launch": (C.c_int, [_vp, _vp]),
    "gf2t_graph_nodes": (C.c_int, [_vp, C.POINTER(_sz)]),
    "gf2t_graph_destroy": (None, [_vp]),
    "gf2t_tune": (C.c_int, [C.c_char_p, C.c_int64]),
    "gf2t_tune_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64)]),
    "gf2t_square_fold": (C.c_int, [_vp, _vp, _u64, _u64, _vp]),
    "gf2t_square_fold_steps": (C.c_int, [_vp, _vp, _u64, _u64, _u64, _vp]),
    "gf2t_square_fold_batch": (C.c_int, [C.POINTER(_vp), C.POINTER(_vp), _p64, _p64, _p64,
                                         C.c_int, _vp]),
    "gf2t_square_fold_slab": (C.c_int, [_vp, _vp, _sz, _u64, _vp, C.c_int, _vp]),
    "gf2t_mulmod_slab_workspace_bytes": (_sz, [_u64, C.c_int]),
    "gf2t_mulmod_slab": (C.c_int, [_vp, _sz, _vp, _sz, C.c_int, _u64, _vp, _vp, _sz, C.c_int,
                                   _vp, _vp]),
    "gf2t_fold": (C.c_int, [_vp, _u64, _vp, _u64, _u64, _vp]),
    "gf2t_mulmod_noalloc": (C.c_int, [_vp, _u64, _vp, _u64, _u64, _u64, _vp, _vp]),
    "gf2t_mulmod_batch_workspace_bytes": (_sz, [C.c_int, _p64, _p64]),
    "gf2t_mulmod_batch": (C.c_int, [C.c_int, C.POINTER(_vp), _p64, C.POINTER(_vp), _p64, _u64,
                                    _p64, C.POINTER(_vp), _vp, _vp]),
    "gf2t_mulx_xor": (C.c_int, [_vp, _vp, _u64, _u64, _vp]),
    "gf2t_set_x": (C.c_int, [_vp, _sz, _vp]),
    "gf2t_set_one": (C.c_int, [_vp, _sz, _vp]),
    "gf2t_xor_x": (C.c_int, [_vp, _vp]),
    "gf2t_cantor_basis": (C.c_int, [_p64, C.POINTER(C.c_int)]),
    "gf2t_fft_plan_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "gf2t_fft_plan_destroy": (C.c_int, [_vp]),
    "gf2t_fft_plan_passes": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gf2t_fft_fwd": (C.c_int, [_vp, _vp, _vp]),
    "gf2t_fft_inv": (C.c_int, [_vp, _vp, _vp]),
    "gf2t_mul_out_words": (_sz, [_u64, _u64]),
    "gf2t_mul_workspace_bytes": (_sz, [_u64, _u64]),
    "gf2t_mul": (C.c_int, [_vp, _u64, _vp, _u64, _vp, _vp, _vp]),
    "gf2t_mulmod": (C.c_int, [_vp, _u64, _vp, _u64, _u64, _u64, _vp, _vp, _vp]),
    "gf2t_sieve": (C.c_int, [_u64, C.c_int, C.c_int, _u64, _vp, C.POINTER(SieveStats)]),
    "gf2t_primitive_poly": (_u64, [C.c_int]),
    "gf2t_scan_run": (C.c_int, [_u64, _p64, _sz, C.POINTER(ScanParams), C.POINTER(_vp)]),
    "gf2t_scan_results_count": (_sz, [_vp]),
    "gf2t_scan_results_get": (C.c_int, [_vp, _sz, C.POINTER(ScanResult), C.c_char_p, _sz]),
    "gf2t_scan_line": (C.c_int, [_vp, _sz, C.c_char_p, _sz]),
    "gf2t_scan_results_free": (None, [_vp]),
    "gf2t_scan_results_times": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), _p64]),
    "gf2t_gcd_host": (C.c_int64, [_p64, _sz, _p64, _sz, _p64, _sz]),
    "gf2t_mul_host": (C.c_int, [_p64, _u64, _p64, _u64, _p64]),
    "gf2t_sieve_release": (C.c_int, []),
    "gf2t_server_create": (C.c_int, [_u64, C.POINTER(ScanParams), _sz, C.POINTER(_vp)]),
    "gf2t_server_capacity": (C.c_int, [_vp, C.POINTER(_sz), C.POINTER(_sz)]),
    "gf2t_server_submit": (C.c_int, [_vp, _u64, _p64, _sz, _u64, _u64]),
    "gf2t_server_step": (C.c_int, [_vp, C.c_int, C.POINTER(_sz), C.POINTER(_sz)]),
    "gf2t_server_poll": (C.c_int, [_vp, _sz, _p64, C.c_char_p, _sz, C.POINTER(_sz)]),
    "gf2t_server_destroy": (None, [_vp]),
    "gf2t_ddf_init_slab": (C.c_int, [_vp, _vp, _sz, C.c_int, _u64, _vp, _u64, C.c_int, _vp]),
    "gf2t_square_fold_slab_g": (C.c_int, [_vp, _vp, _sz, _u64, _vp, _sz, C.c_int, _vp]),
    "gf2t_square2_fold_slab_g": (C.c_int, [_vp, _vp, _sz, _u64, _vp, _sz, C.c_int, _vp]),
    "gf2t_sq2_slab_plan_create": (C.c_int, [_u64, _p64, _sz, C.POINTER(_vp)]),
    "gf2t_sq2_slab_plan_set": (C.c_int, [_vp, _p64, _sz, _vp]),
    "gf2t_sq2_slab_plan_destroy": (C.c_int, [_vp]),
    "gf2t_square2_fold_slab_p": (C.c_int, [_vp, _vp, _vp, _sz, _sz, C.c_int, _vp]),
    "gf2t_horner_slab": (C.c_int, [_vp, _sz, C.c_int, _u64, _vp, _vp, _sz, C.c_int, _vp]),
    "gf2t_mulx_xor_slab": (C.c_int, [_vp, _sz, _vp, _sz, _u64, _vp, C.c_int, _vp]),
    "gf2t_xor_word0_slab": (C.c_int, [_vp, _sz, _u64, C.c_int, _vp]),
    "gf2t_set_one_slab": (C.c_int, [_vp, _sz, _sz, C.c_int, _vp]),
    "gf2t_copy_slab": (C.c_int, [_vp, _sz, _vp, _sz, _sz, C.c_int, _vp]),
    "gf2t_mat2_host": (C.c_int, [C.POINTER(_p64), _p64, _p64, _u64, _p64, _u64, _p64, _p64, _sz]),
    "gf2t_mat2n_host": (C.c_int, [C.POINTER(_p64), _p64, C.c_int, C.POINTER(_p64), _p64,
                                  C.POINTER(_p64), _sz]),
    "gf2t_edf_least": (C.c_int, [_p64, _sz, _u64, _p64, _sz]),
    "gf2t_swan_bound": (_u64, [_u64, _u64, C.POINTER(C.c_int)]),
}


def declared_symbols() -> list[str]:
    """Every function name declared in include/gf2t.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(gf2t_[a-z0-9_]+)\s*\(", txt)))


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP engine library is required; there is no other compute path)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib.gf2t_last_error().decode(errors="replace")
        raise GF2TError(rc, f"{what}: {msg}" if what else msg)
