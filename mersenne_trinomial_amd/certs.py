"""Certificate and worklist formats of the coarse sieve and the deep scan.

  's d p<hex>'  least factor: degree d, lowercase hex bit mask of d+1 bits,
                MSB first, no leading zero; ties go to the numerically
                least mask (README.md:79-104, src/sieve_ref.cpp:163-169)
  's u'         search bound reached without a factor
  's primitive' no factor up to the Swan bound
  's rA-B'      a factor with degree in [A, B], unresolved
  <prefix>.found.cert     sieve lines, ascending s, survivors and 0
                          entries skipped (src/coarse_sieve.cu:1528-1549)
  <prefix>.survivors.txt  one s per line (src/coarse_sieve.cu:1510-1526)
  <prefix>.found.bin      little-endian u64 keys (d<<48)|mask for
                          s = 1..smax, ~0 = survivor (src/coarse_sieve.cu:
                          1551-1558)
"""
from __future__ import annotations

import os

import numpy as np

NONE = np.uint64(0xFFFFFFFFFFFFFFFF)
MASK48 = np.uint64(0xFFFFFFFFFFFF)


def key_line(s: int, key: int) -> str:
    return f"{s} {key >> 48} p{key & 0xFFFFFFFFFFFF:x}"


def sieve_lines(best: np.ndarray):
    """Certificate lines for every s >= 1 with a factor (key != ~0, != 0)."""
    best = np.asarray(best, dtype=np.uint64)
    idx = np.nonzero((best != NONE) & (best != 0))[0]
    idx = idx[idx >= 1]
    d = (best[idx] >> np.uint64(48)).astype(np.int64)
    m = (best[idx] & MASK48)
    return [f"{s} {dd} p{int(mm):x}" for s, dd, mm in zip(idx.tolist(), d.tolist(), m.tolist())]


def survivors(best: np.ndarray) -> np.ndarray:
    best = np.asarray(best, dtype=np.uint64)
    s = np.nonzero(best == NONE)[0]
    return s[s >= 1]


def write_sieve_outputs(prefix: str, best: np.ndarray, cert: bool = True,
                        surv: bool = True, found_bin: bool = False) -> dict:
    """Write <prefix>.found.cert / .survivors.txt / .found.bin."""
    out = {}
    smax = len(best) - 1
    if cert:
        p = prefix + ".found.cert"
        with open(p, "w") as f:
            for ln in sieve_lines(best):
                f.write(ln + "\n")
        out["cert"] = p
    if surv:
        p = prefix + ".survivors.txt"
        np.savetxt(p, survivors(best), fmt="%d")
        out["survivors"] = p
    if found_bin:
        p = prefix + ".found.bin"
        np.asarray(best[1:], dtype="<u8").tofile(p)
        out["found_bin"] = p
    out["smax"] = smax
    return out


def read_found_bin(path: str, smax: int) -> np.ndarray:
    """best[0..smax] from a .found.bin image (entry 0 unused, ~0)."""
    data = np.fromfile(path, dtype="<u8")
    if data.size != smax:
        raise ValueError(f"{path}: {data.size} entries, expected {smax}")
    best = np.empty(smax + 1, dtype=np.uint64)
    best[0] = NONE
    best[1:] = data
    return best


def read_survivors(path: str) -> np.ndarray:
    if os.path.getsize(path) == 0:
        return np.zeros(0, dtype=np.uint64)
    return np.loadtxt(path, dtype=np.uint64, ndmin=1)


def parse_line(line: str):
    """(s, kind, d, hex) for a certificate line."""
    parts = line.split()
    s = int(parts[0])
    if parts[1] == "u":
        return s, "u", None, None
    if parts[1] == "primitive":
        return s, "primitive", None, None
    if parts[1].startswith("r"):
        return s, "range", None, parts[1][1:]
    return s, "p", int(parts[1]), parts[2][1:]


def best_from_survivors(surv, smax: int, min_depth: int) -> np.ndarray:
    """A best[] image seeded from a survivors list (src/coarse_sieve.cu:
    1449-1477): every s not listed gets a sentinel below min_depth's keys,
    every listed s stays ~0."""
    best = np.full(smax + 1, np.uint64((min_depth - 1) << 48), dtype=np.uint64)
    best[0] = NONE
    surv = np.asarray(surv, dtype=np.int64)
    best[surv[(surv >= 1) & (surv <= smax)]] = NONE
    return best
