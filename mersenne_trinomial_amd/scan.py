"""Deep scan of sieve survivors over the C ABI (gf2t_scan_run).

Mirrors trinomial_scan / tsfactor: for each s, the least factor degree d
in (k0, bound] and the numerically least degree-d factor, as certificate
lines 's d p<hex>' / 's u' / 's primitive' / 's rA-B'
(src/tsfactor.cu:2480-2509).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._capi import ScanParams, ScanResult, check, lib
from .device import init

FOUND, GAVE_UP, PRIMITIVE, RANGE = 1, 2, 3, 4


@dataclass
class ScanOut:
    s: int
    kind: int
    d: int = 0
    hex: str = ""
    rlo: int = 0
    rhi: int = 0
    gcd_degree: int = 0

    def line(self) -> str:
        if self.kind == FOUND:
            return f"{self.s} {self.d} p{self.hex}"
        if self.kind == GAVE_UP:
            return f"{self.s} u"
        if self.kind == PRIMITIVE:
            return f"{self.s} primitive"
        return f"{self.s} r{self.rlo}-{self.rhi}"


def swan_bound(r: int, s: int):
    sw = C.c_int()
    b = lib.gf2t_swan_bound(r, s, C.byref(sw))
    return int(b), bool(sw.value)


def scan(r: int, svals, k0: int, maxd: int = 0, L0: int = 64, Lmax: int = 4096,
         gcd_threads: int = 0, batch: int = 0, verbose: int = 0,
         block_m: int = 0, past_bound: bool = False, times: dict | None = None) -> list[ScanOut]:
    """Scan every s in svals (sieve survivors at depth k0).  past_bound:
    report a factor found in the last whole block beyond the degree bound
    (tsfactor's semantics) instead of 'u' / 'primitive'.  batch = 1 runs
    the s values one at a time (the reference's per-s loop); 0 puts as
    many in one device cohort as memory allows.  times (a dict) receives
    the run's phase seconds: device_s, gcd_s, resolve_s, and n_gcd."""
    init()
    sv = np.ascontiguousarray(np.asarray(svals, dtype=np.uint64))
    P = ScanParams(k0=k0, maxd=maxd, L0=L0, Lmax=Lmax, gcd_threads=gcd_threads, batch=batch,
                   verbose=verbose, block_m=block_m, past_bound=int(bool(past_bound)))
    res = C.c_void_p()
    check(lib.gf2t_scan_run(r, sv.ctypes.data_as(C.POINTER(C.c_uint64)), sv.size, C.byref(P),
                            C.byref(res)), "gf2t_scan_run")
    try:
        if times is not None:
            tv = [C.c_double() for _ in range(3)]
            ng = C.c_uint64()
            check(lib.gf2t_scan_results_times(res, *[C.byref(x) for x in tv], C.byref(ng)))
            times.update(device_s=tv[0].value, gcd_s=tv[1].value, resolve_s=tv[2].value,
                         n_gcd=int(ng.value))
        out = []
        n = lib.gf2t_scan_results_count(res)
        for i in range(n):
            row = ScanResult()
            check(lib.gf2t_scan_results_get(res, i, C.byref(row), None, 0))
            hx = ""
            if row.hex_len:
                buf = C.create_string_buffer(row.hex_len + 1)
                check(lib.gf2t_scan_results_get(res, i, None, buf, row.hex_len + 1))
                hx = buf.value.decode()
            out.append(ScanOut(int(row.s), int(row.kind), int(row.d), hx, int(row.rlo),
                               int(row.rhi), int(row.gcd_degree)))
        return out
    finally:
        lib.gf2t_scan_results_free(res)
