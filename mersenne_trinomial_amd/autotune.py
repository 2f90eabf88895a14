"""Autotuner: measures the engine's tuning knobs on this GPU and writes the
winners to gf2t_tune.txt beside libgf2t.so, which gf2t_init applies in
every later process (CLIs, Python, bench.py).

What it searches (the reference times its transform variants and picks
its block size the same way, src/tsfactor.cu:1986-1989 and auto_m
:2093-2114):

- fft_tile_<m>: the tile of the Cantor transform's LDS passes, 2^T
  elements for T in 6..13, per transform size 2^m.  Small tiles spread a
  small transform over more workgroups but need more passes.  Timed on
  the product shape the half-gcd's device service runs most: a batch of
  six transforms (four matrix entries and two vectors), here as three
  two-operand products back to back on one stream.
- the batched squaring knobs sq_fuse / sq_nt / sq_order / sq_ppt, timed
  on the deep scan's residue shape at the given r.
- gcd_dev_bits: the product size from which the half-gcd uses the device,
  timed on concurrent full-size gcds (the deep scan's interval gcds at
  r = 136279841).
- gcd_base_bits: the half-gcd's base-case size, on the same concurrent gcds.
- block_m_<k>: the deep scan's degrees per interval product for
  2^k <= r < 2^(k+1), from the measured per-degree cost
  m * t_sq + t_mul / m of the blocked scheme (t_sq and t_mul: a
  residue's share of a cohort's slab squaring and slab product).

    python -m mersenne_trinomial_amd.autotune [--out PATH] [--quick]

Prints one JSON line with every knob's default and chosen value and
their times.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from ._capi import LIB_PATH, check, lib
from .device import DeviceBuffer, Event, init, synchronize

P64 = C.POINTER(C.c_uint64)


def tune_get(key: str) -> int:
    v = C.c_int64()
    check(lib.gf2t_tune_get(key.encode(), C.byref(v)), f"gf2t_tune_get {key}")
    return int(v.value)


def tune_set(key: str, value: int) -> None:
    check(lib.gf2t_tune(key.encode(), int(value)), f"gf2t_tune {key}")


def _time_ms(fn, reps: int) -> float:
    fn()
    synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    synchronize()
    return e0.elapsed_ms(e1) / reps


def tune_fft(ms, tiles, reps: int = 5):
    """For each transform size 2^m, the tile bits with the least time for
    three products whose transforms have 2^m points."""
    rng = np.random.default_rng(1)
    out = {}
    for m in ms:
        bits = 16 * (1 << m) - 40            # two such operands need 2^m points
        words = (bits + 63) // 64
        a = DeviceBuffer.from_array(rng.integers(0, 2**64, words, dtype=np.uint64))
        b = DeviceBuffer.from_array(rng.integers(0, 2**64, words, dtype=np.uint64))
        dout = DeviceBuffer.words(lib.gf2t_mul_out_words(bits, bits))
        work = DeviceBuffer(lib.gf2t_mul_workspace_bytes(bits, bits))

        def run():
            for _ in range(3):
                check(lib.gf2t_mul(a.ptr, bits, b.ptr, bits, dout.ptr, work.ptr, None))

        default = tune_get(f"fft_tile_{m}")
        times = {}
        for T in tiles:
            if T > max(m, 6) or T < 6:
                continue
            tune_set(f"fft_tile_{m}", T)
            times[T] = _time_ms(run, reps)
        best = min(times, key=times.get)
        tune_set(f"fft_tile_{m}", best)
        out[m] = {"default": default, "default_ms": times.get(default), "chosen": best,
                  "chosen_ms": times[best], "ms": {T: round(v, 4) for T, v in times.items()}}
        print(f"fft m={m}: tile 2^{best} {times[best]:.3f} ms (default 2^{default} "
              f"{times.get(default, float('nan')):.3f} ms)", file=sys.stderr, flush=True)
    return out


def tune_squaring(r: int, n: int, steps: int, reps: int = 2):
    """The batched squaring knobs on n residues of r bits (steps squarings
    per chain), one knob at a time from the current settings."""
    nw = (r + 63) // 64
    res = [DeviceBuffer.words(nw) for _ in range(n)]
    tmp = [DeviceBuffer.words(nw) for _ in range(n)]
    for x in res:
        check(lib.gf2t_set_x(x.ptr, nw, None))
    rng = np.random.default_rng(2)
    svals = np.sort(rng.choice(np.arange(3, (r - 1) // 2), n, replace=False)).astype(np.uint64)
    rr = np.full(n, r, dtype=np.uint64)
    kk = np.full(n, steps, dtype=np.uint64)
    cres = (C.c_void_p * n)(*[x.ptr for x in res])
    ctmp = (C.c_void_p * n)(*[x.ptr for x in tmp])

    def run():
        check(lib.gf2t_square_fold_batch(cres, ctmp, rr.ctypes.data_as(P64),
                                         svals.ctypes.data_as(P64), kk.ctypes.data_as(P64), n,
                                         None))

    out = {}
    knobs = [("sq_fuse", (0, 1)), ("sq_nt", (0, 1)), ("sq_order", (0, 1)), ("sq_ppt", (1, 2, 4))]
    for key, values in knobs:
        if key in ("sq_order", "sq_ppt") and out["sq_fuse"]["chosen"] == 1:
            continue    # knobs of the one-squaring kernel only
        default = tune_get(key)
        times = {}
        for v in values:
            tune_set(key, v)
            times[v] = _time_ms(run, reps)
        best = min(times, key=times.get)
        tune_set(key, best)
        out[key] = {"default": default, "default_ms": round(times[default], 3),
                    "chosen": best, "chosen_ms": round(times[best], 3)}
        print(f"{key}: {best} ({times[best]:.3f} ms; default {default} {times[default]:.3f} ms)",
              file=sys.stderr, flush=True)
    return out


def tune_block_m(r: int, reps: int = 3):
    """Degrees per interval product for r's range, from the blocked
    scheme's per-degree cost m t_sq + t_mul / m, with t_sq and t_mul the
    per-residue times of a cohort's slab squaring and slab product."""
    nw = (r + 63) // 64
    pitch = (nw + 31) & ~31
    n = 8 if r > (1 << 24) else 64
    a = DeviceBuffer.words(n * pitch)
    b = DeviceBuffer.words(n * pitch)
    o = DeviceBuffer.words(n * pitch)
    for i in range(n):                          # valid residues (bits >= r clear)
        check(lib.gf2t_set_x(a.word_ptr(i * pitch), nw, None))
        check(lib.gf2t_set_x(b.word_ptr(i * pitch), nw, None))
    hs = np.array([(r - 1) // 3 - i for i in range(n)], dtype=np.uint64)
    s = DeviceBuffer.from_array(hs)
    work = DeviceBuffer(lib.gf2t_mulmod_slab_workspace_bytes(r, n))
    # per squaring, in the launches the deep scan uses for this r: two
    # squarings per launch in work units from 2^18 bits (when sq_fuse is
    # on), else one
    if r >= (1 << 18) and tune_get("sq_fuse"):
        plan = C.c_void_p()
        check(lib.gf2t_sq2_slab_plan_create(r, hs.ctypes.data_as(P64), n, C.byref(plan)))
        try:
            t_sq = _time_ms(lambda: check(lib.gf2t_square2_fold_slab_p(plan, a.ptr, o.ptr, pitch,
                                                                        n, 1, None)),
                            8 * reps) / (2 * n)
        finally:
            synchronize()
            check(lib.gf2t_sq2_slab_plan_destroy(plan))
    else:
        t_sq = _time_ms(lambda: check(lib.gf2t_square_fold_slab_g(a.ptr, o.ptr, pitch, r, s.ptr,
                                                                   n, 1, None)), 8 * reps) / n
    t_mul = _time_ms(lambda: check(lib.gf2t_mulmod_slab(a.ptr, pitch, b.ptr, pitch, 0, r, s.ptr,
                                                        o.ptr, pitch, n, work.ptr, None)),
                     reps) / n
    cost = {m: m * t_sq + t_mul / m for m in (4, 8, 12, 16, 20, 24, 32, 40, 48, 56, 64)
            if 2 * m < r}
    best = min(cost, key=cost.get)
    k = int(math.floor(math.log2(r)))
    default = tune_get(f"block_m_{k}") or (4 if r < (1 << 14) else 16 if r < (1 << 19) else 32)
    tune_set(f"block_m_{k}", best)
    print(f"block_m r={r}: {best} (t_sq {t_sq:.4f} ms, t_mul {t_mul:.3f} ms, "
          f"{cost[best]:.4f} ms/degree)", file=sys.stderr, flush=True)
    return {"r": r, "k": k, "t_sq_ms": t_sq, "t_mul_ms": t_mul, "default": default,
            "chosen": best, "ms_per_degree": {m: round(v, 4) for m, v in cost.items()}}


def tune_gcd(r: int, conc: int, thresholds, key: str = "gcd_dev_bits"):
    """gcd_dev_bits (or another gcd knob named by `key`, such as
    gcd_base_bits): the value among `thresholds` that runs `conc`
    concurrent gcds of random r-bit residues against x^r + x^3 + 1
    fastest (the deep scan runs its interval gcds concurrently).
    gcd_dev_bits is the size from which the half-gcd's products go to the
    device; gcd_base_bits the size below which its base case runs."""
    import threading
    nw = (r + 63) // 64
    t = np.zeros((r + 64) // 64, dtype=np.uint64)
    for e in (0, 3, r):
        t[e // 64] |= np.uint64(1 << (e % 64))
    rng = np.random.default_rng(3)
    As = []
    for _ in range(conc):
        a = rng.integers(0, 2**64, nw, dtype=np.uint64)
        if r % 64:
            a[-1] &= np.uint64((1 << (r % 64)) - 1)
        As.append(a)

    def run_all():
        def one(a):
            g = np.zeros(4, dtype=np.uint64)
            lib.gf2t_gcd_host(t.ctypes.data_as(P64), t.size, a.ctypes.data_as(P64), a.size,
                              g.ctypes.data_as(P64), 4)
        ths = [threading.Thread(target=one, args=(a,)) for a in As]
        t0 = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        return time.perf_counter() - t0

    default = tune_get(key)
    times = {}
    for b in thresholds:
        tune_set(key, b)
        run_all()                                   # plans and buffers for this size mix
        times[b] = run_all()
    best = min(times, key=times.get)
    tune_set(key, best)
    print(f"{key}: {best} ({times[best]:.3f} s for {conc} gcds at r={r}; default "
          f"{default} {times.get(default, float('nan')):.3f} s)", file=sys.stderr, flush=True)
    return {"key": key, "r": r, "concurrent": conc, "default": default, "chosen": best,
            "s": {b: round(v, 3) for b, v in times.items()}}


def write(path: str, fft: dict, sq: dict, blocks: list, device: str, gcd: dict | None,
          gcd_base: dict | None = None) -> None:
    lines = [f"# gf2t autotuner ({device.strip()}), {time.strftime('%Y-%m-%d %H:%M:%S')}",
             "# key value -- applied by gf2t_init; see mersenne_trinomial_amd/autotune.py"]
    for m, v in sorted(fft.items()):
        lines.append(f"fft_tile_{m} {v['chosen']}")
    for key, v in sq.items():
        lines.append(f"{key} {v['chosen']}")
    for b in blocks:
        lines.append(f"block_m_{b['k']} {b['chosen']}")
    if gcd:
        lines.append(f"gcd_dev_bits {gcd['chosen']}")
    if gcd_base:
        lines.append(f"gcd_base_bits {gcd_base['chosen']}")
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(tmp, path)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(LIB_PATH), "gf2t_tune.txt"))
    ap.add_argument("--quick", action="store_true", help="fewer sizes and repetitions")
    ap.add_argument("--sq-r", type=int, default=136279841)
    ap.add_argument("--sq-residues", type=int, default=64)
    ap.add_argument("--block-r", type=str, default="132049,756839,3021377,136279841")
    ap.add_argument("--gcd-r", type=int, default=136279841)
    ap.add_argument("--gcd-conc", type=int, default=15,
                    help="gcds at once when timing the gcd threshold (the deep scan's survivors "
                         "of 256 uniformly drawn s at r = 136279841 number about 15); 0 skips it")
    args = ap.parse_args(argv)
    os.environ["GF2T_TUNE_FILE"] = ""          # measure from the built-in defaults
    init(0)
    from .device import device_name
    dev = device_name()
    ms = range(10, 25, 2) if args.quick else range(8, 25)
    t0 = time.perf_counter()
    fft = tune_fft(ms, range(6, 14), reps=3 if args.quick else 5)
    sq = tune_squaring(args.sq_r, args.sq_residues, 8 if args.quick else 16)
    blocks = [tune_block_m(int(r)) for r in args.block_r.split(",") if r]
    gcd = tune_gcd(args.gcd_r, args.gcd_conc,
                   [1 << k for k in range(15, 19)]) if args.gcd_conc > 0 else None
    # the base-case size, timed with the chosen device threshold
    gcd_base = tune_gcd(args.gcd_r, args.gcd_conc, [1536, 2048, 3072, 4096],
                        key="gcd_base_bits") if args.gcd_conc > 0 else None
    write(args.out, fft, sq, blocks, dev, gcd, gcd_base)
    print(json.dumps({"mode": "autotune", "device": dev.strip(), "file": args.out,
                      "seconds": round(time.perf_counter() - t0, 1), "fft_tile": fft,
                      "squaring": sq, "block_m": blocks, "gcd": gcd, "gcd_base": gcd_base}),
          flush=True)


if __name__ == "__main__":
    main()
