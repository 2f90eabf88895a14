"""Benchmark: the GF(2) trinomial hot path at r = 136279841 on one MI355X.

One step = one pass of the GPU hot path over one batch of synthetic input:
  (a) Zech-log coarse sieve, degrees 2..32, every s in [1, 68139920]:
      each s with an irreducible factor of degree <= 32 gets its final
      certificate line 's d p<hex>' (smallest degree, least mask);
  (b) Frobenius chains F <- F^2 mod x^r + x^s + 1 for 256 sieve survivors
      (seeded draw; 17 MB residues resident in HBM, 4.36 GB in flight)
      advanced 512 squarings (one checkpoint interval) in the batched
      form, starting from F = x; two squarings per launch
      (k_square2_fold_batch: F is read once and F^4 written once).
roofline = the squaring kernel of the chain: algorithmic bytes 16*nw per
residue per launch (read the residue, write its square or fourth power)
times 256 residues, over its average launch time measured with HIP
events on its stream.

value = s-values certified per second on the deep-scan workload, one
pass timed after the steps: 256 s drawn by SplitMix64 from
[3, 68139920]; the coarse sieve to degree 32 certifies most of them, and
its survivors go through the batched deep scan (blocked distinct-degree
products on the device, an interval gcd every 512 degrees on host
threads with device products, resolution of hits) to degree 4096, each s
leaving at its first hit.  The same s values then run through the plain
form (one s at a time, as the reference's per-s loop runs them) and
their certificate lines must equal the batched ones (`c3` object).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

R = 136279841
SMAX = R // 2
HBM_PEAK_GBS = 8000.0
# HBM bytes one launch of each squaring kernel moves for 256 residues at
# r = 136279841, from rocprofv3 FETCH_SIZE and WRITE_SIZE runs scaled by
# the factors a copy kernel of known size gives (x2.0 for FETCH_SIZE, x1.0
# for WRITE_SIZE).  The command lines and raw counts are kept in
# profiles/k_square_fold_batch_r136279841_256res.md and
# profiles/k_square2_fold_batch_r136279841_256res.md.
SQ_TRAFFIC_256 = {
    "k_square_fold_batch": 6732439347 + 4360962560,    # one squaring per launch
    "k_square2_fold_batch": 8157166592 + 4366682419,   # two squarings per launch
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--depth", type=int, default=32)
    ap.add_argument("--residues", type=int, default=256)
    ap.add_argument("--chain-steps", type=int, default=512)
    ap.add_argument("--seed", type=int, default=20250101)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sq-ppt", type=int, default=0, help="override the squaring pairs/lane")
    ap.add_argument("--sq-order", type=int, default=-1, help="override the squaring tile order")
    ap.add_argument("--sq-nt", type=int, default=-1, help="override streaming stores (0/1)")
    ap.add_argument("--sq-fuse", type=int, default=-1,
                    help="override two squarings per launch (0/1; default on)")
    ap.add_argument("--scan-r", type=int, default=R, help="scan mode: trinomial degree r")
    ap.add_argument("--scan-s", type=str, default="474",
                    help="scan mode: comma-separated s values, or survivors:N for the first N "
                         "sieve survivors at depth --scan-k0")
    ap.add_argument("--scan-k0", type=int, default=32, help="scan mode: depth already sieved")
    ap.add_argument("--scan-maxd", type=int, default=64, help="scan mode: degree bound")
    ap.add_argument("--scan-l0", type=int, default=64, help="scan mode: first interval length")
    ap.add_argument("--scan-block", type=str, default="16,1",
                    help="scan mode: comma-separated block sizes m to run (1 = plain)")
    ap.add_argument("--out", type=str, default="",
                    help="pipeline mode: write the certificate lines to this file")
    ap.add_argument("--serve-jobs", type=int, default=10, help="serve mode: number of jobs")
    ap.add_argument("--serve-rate", type=float, default=0.5,
                    help="serve mode: Poisson arrival rate, jobs per second")
    ap.add_argument("--serve-max-len", type=int, default=5000,
                    help="serve mode: longest s slice (the mix draws 10^3..10^6, capped here)")
    ap.add_argument("--serve-interval", type=int, default=2048,
                    help="serve mode: deep-scan interval length (gcds per s = 2048 / this)")
    ap.add_argument("--c3", type=int, default=1,
                    help="run the deep-scan workload (256 seeded s, sieve then deep scan) once "
                         "after the timed steps; its rate is the line's value")
    ap.add_argument("--c4", type=int, default=1,
                    help="run the Cantor-product batch (64 seeded pairs of 2^16..2^27 bits, each "
                         "folded) after the timed steps and report it in a c4 object")
    ap.add_argument("--c3-n", type=int, default=256)
    ap.add_argument("--c3-seed", type=int, default=20250101)
    ap.add_argument("--c3-maxd", type=int, default=4096)
    ap.add_argument("--c3-interval", type=int, default=512)
    ap.add_argument("--c3-plain", type=int, default=1,
                    help="also run the same s one at a time (the plain per-s form)")
    ap.add_argument("--scan-stagger", type=int, default=None,
                    help="parts a deep-scan cohort is staggered into (gf2t_tune scan_stagger: "
                         "0 = automatic, -1 = never)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="a gf2t_tune setting applied after the tuning file (repeatable)")
    ap.add_argument("--gcd-conc", type=int, default=8,
                    help="gcd mode: gcds run at once after the single one")
    ap.add_argument("--mode",
                    choices=["full", "squaring", "fft", "scan", "pipeline", "serve", "sieve",
                             "c3", "gcd", "fft24"],
                    default="full",
                    help="full: the default bench line; squaring: a plain copy of known size "
                         "then only the batched squaring chain (for rocprofv3 counter runs); "
                         "fft: the Cantor-product batch (64 seeded operand pairs of 2^16..2^27 "
                         "bits, each product folded mod x^r + x^s + 1); sieve: --steps full-range "
                         "coarse sieves at r = 136279841 to --depth (up to 36), per-degree "
                         "device times")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(certified: int, chain_s: int, c3: dict | None):
    """The oracle (the CPU restatement) on the same inputs as the sieve leg:
    r = 136279841, every s <= 68139920, degrees 2..32.  On all the job's
    cores: degrees 2..26 in one run, then 27, and 28..32 extrapolated from
    degree 27 (per-degree work is dominated by the 2^d
    classes and table entries, so each degree costs twice the one before).
    On one core: 2..22, then 23 and 24, and 25..32 extrapolated from 24.  Plus
    two full-size squaring+fold steps on one core, and from those a lower
    bound on the CPU time of the deep-scan workload (its squarings only)."""
    import oracle as O
    cores = host_threads()

    def sieve_time(nthreads, d_all, d_each):
        t0 = time.perf_counter()
        best, _ = O.sieve(R, d_all, SMAX, nthreads=nthreads)
        t_all = time.perf_counter() - t0
        per = {}
        for d in d_each:
            t0 = time.perf_counter()
            best, _ = O.sieve(R, d, SMAX, dmin=d, nthreads=nthreads, best=best)
            per[d] = time.perf_counter() - t0
        del best
        last = d_each[-1]
        extra = {d: per[last] * 2.0 ** (d - last) for d in range(last + 1, 33)}
        total = t_all + sum(per.values()) + sum(extra.values())
        return total, {"measured_2_to_%d_s" % d_all: round(t_all, 3),
                       "measured_per_degree_s": {d: round(v, 3) for d, v in per.items()},
                       "extrapolated_per_degree_s": {d: round(v, 2) for d, v in extra.items()},
                       "total_s": round(total, 2)}

    t_n, info_n = sieve_time(cores, 26, [27])
    t_1, info_1 = sieve_time(1, 22, [23, 24])
    x = O.x_residue(R)
    t0 = time.perf_counter()
    O.square_fold_steps(x, R, chain_s, 2)
    t_sq = (time.perf_counter() - t0) / 2
    out = {
        "value": certified / t_n,
        "unit": "s-values certified/sec",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle coarse sieve at r={R} over all s<=68139920, degrees 2..32 (the "
                   f"sieve leg's inputs; {certified} certified): degrees 2..27 measured on "
                   f"{cores} threads, 28..32 extrapolated (x2 per degree) = {t_n:.1f} s; "
                   f"one core: {t_1:.1f} s (2..24 measured)"),
        "nproc": os.cpu_count(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
        "cpu_model": cpu_model(),
        "sieve_all_cores": info_n,
        "sieve_one_core": info_1,
        "one_core_value": certified / t_1,
        "squaring_fold_residue_steps_per_s_1core": 1.0 / t_sq,
    }
    out["sieve_s_per_s"] = certified / t_n
    if c3:
        # the deep-scan workload: the same sieve, then for its survivors the
        # squarings of the first interval (degrees 33..544) of the blocked
        # scheme, 32 per degree, spread over the cores; the interval
        # products and gcds are left out, so this bounds the CPU time from
        # below and its rate from above
        sq = c3["survivors"] * 512 * 32 * t_sq / cores
        out["value"] = c3["n_s"] / (t_n + sq)
        out["sample"] = (f"deep-scan workload ({c3['n_s']} s) on the oracle, {cores} threads: the "
                         f"coarse sieve of all s<=68139920 to degree 32 ({t_n:.1f} s: degrees "
                         f"2..27 measured, 28..32 extrapolated x2 per degree) + the "
                         f"{c3['survivors']} survivors' first-interval squarings "
                         f"({c3['survivors']} x 512 x 32 x {1e3 * t_sq:.1f} ms / {cores} = "
                         f"{sq:.1f} s); interval products and gcds omitted, so value is an "
                         f"upper bound on the CPU rate.  Sieve leg alone: "
                         f"{certified / t_n:.3g} s/s on {cores} threads, "
                         f"{certified / t_1:.3g} s/s on one core")
        out["c3_cpu_lower_bound_s"] = round(t_n + sq, 1)
    return out


M64 = (1 << 64) - 1


def splitmix_draw(seed: int, n: int, lo: int, hi: int) -> list[int]:
    """n distinct integers in [lo, hi] from SplitMix64 (Steele, Lea and
    Flood; the generator the reference seeds its random inputs with,
    src/tsfactor.cu:1993-1998), in draw order."""
    st, out, seen = seed & M64, [], set()
    while len(out) < n:
        st = (st + 0x9E3779B97F4A7C15) & M64
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        z ^= z >> 31
        v = lo + z % (hi - lo + 1)
        if v not in seen:
            seen.add(v)
            out.append(v)
    return out


def c3_svals(seed: int, n: int = 256) -> list[int]:
    """The deep-scan workload's s values: n distinct s drawn uniformly from
    [3, 68139920] by SplitMix64 from `seed`."""
    return splitmix_draw(seed, n, 3, SMAX)


def c3_run(svals, depth: int, maxd: int, interval: int, batch: int, gcd_threads: int,
           dbest=None):
    """Certify every s of svals at r = 136279841: the coarse sieve over all
    s <= 68139920 to `depth`, its lines for the s it certifies, then the
    deep scan of the rest from degree depth+1 to maxd with an interval gcd
    every `interval` degrees and per-s early exit.  batch = 0: one device
    cohort (the batched form); batch = 1: one s at a time (the reference's
    per-s loop).  Returns (lines by s, phase times)."""
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    own = dbest is None
    if own:
        dbest = DeviceBuffer.words(SMAX + 1)
    t0 = time.perf_counter()
    dbest.fill(0xFF)
    check(lib.gf2t_sieve(R, 2, depth, SMAX, dbest.ptr, None), "gf2t_sieve")
    key = np.zeros(1, dtype=np.uint64)
    lines, surv = {}, []
    for s in svals:
        check(lib.gf2t_memcpy_d2h(key.ctypes.data, dbest.ptr + 8 * s, 8, None))
        k = int(key[0])
        if k == 0xFFFFFFFFFFFFFFFF:
            surv.append(s)
        else:
            lines[s] = f"{s} {k >> 48} p{k & 0xFFFFFFFFFFFF:x}"
    device.synchronize()
    t1 = time.perf_counter()
    times = {}
    out = M.scan(R, surv, k0=depth, maxd=maxd, L0=interval, Lmax=interval, block_m=0,
                 gcd_threads=gcd_threads, batch=batch, times=times, verbose=1) if surv else []
    t2 = time.perf_counter()
    kinds = {}
    for o in out:
        lines[o.s] = o.line()
        kinds[o.kind] = kinds.get(o.kind, 0) + 1
    if own:
        del dbest
    return lines, {"seconds": t2 - t0, "sieve_s": t1 - t0, "scan_s": t2 - t1,
                   "survivors": len(surv), "found": kinds.get(1, 0), "u": kinds.get(2, 0),
                   **{k: (round(v, 3) if isinstance(v, float) else v) for k, v in times.items()}}


def c3_measure(args, dbest):
    """The batched deep scan and the plain one-s-at-a-time form on the same
    seeded s values; the certificate lines must be identical."""
    svals = c3_svals(args.c3_seed, args.c3_n)
    threads = host_threads()
    kw = dict(depth=args.depth, maxd=args.c3_maxd, interval=args.c3_interval,
              gcd_threads=threads, dbest=dbest)
    print(f"c3: {len(svals)} s (seed {args.c3_seed}), batched form", file=sys.stderr, flush=True)
    lines_b, tb = c3_run(svals, batch=0, **kw)
    print(f"c3: batched {tb}", file=sys.stderr, flush=True)
    res = {"n_s": len(svals), "seed": args.c3_seed, "depth": args.depth, "maxd": args.c3_maxd,
           "interval": args.c3_interval, "gcd_threads": threads,
           "s_per_s": len(svals) / tb["seconds"], **tb}
    if args.c3_plain:
        print("c3: plain form (one s at a time)", file=sys.stderr, flush=True)
        lines_p, tp = c3_run(svals, batch=1, **kw)
        print(f"c3: plain {tp}", file=sys.stderr, flush=True)
        res["plain"] = {"s_per_s": len(svals) / tp["seconds"], **tp}
        res["ratio"] = tp["seconds"] / tb["seconds"]
        res["certificates_equal"] = lines_p == lines_b
    import hashlib
    text = "".join(lines_b[s] + "\n" for s in sorted(lines_b))
    res["lines_md5"] = hashlib.md5(text.encode()).hexdigest()
    res["lines_sample"] = [lines_b[s] for s in sorted(lines_b)
                           if " p" in lines_b[s] and int(lines_b[s].split()[1]) > args.depth][:8]
    return res


def host_threads() -> int:
    """Host threads this process is given: OMP_NUM_THREADS when set (the
    GPU box sets it to the job's CPU share), else the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    aff = len(os.sched_getaffinity(0))
    return min(aff, int(env)) if env.isdigit() and int(env) > 0 else aff


def mode_squaring(args):
    """Copy of 16*nw*residues bytes (calibrates rocprofv3 FETCH/WRITE_SIZE
    against a kernel of known traffic), then the batched chain alone."""
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer, Event
    device.init(0)
    nw = (R + 63) // 64
    n = args.residues
    res = [DeviceBuffer.words(nw) for _ in range(n)]
    tmp = [DeviceBuffer.words(nw) for _ in range(n)]
    big_a = DeviceBuffer.words(nw * n)
    big_b = DeviceBuffer.words(nw * n)
    e0, e1 = Event(), Event()
    nbytes = 8 * nw * n
    nbytes16 = nbytes & ~15
    check(lib.gf2t_stream_copy(big_b.ptr, big_a.ptr, nbytes16, None))
    device.synchronize()
    e0.record()
    check(lib.gf2t_stream_copy(big_b.ptr, big_a.ptr, nbytes16, None))
    e1.record()
    copy_ms = e0.elapsed_ms(e1)
    e0.record()
    check(lib.gf2t_memcpy_d2d(big_b.ptr, big_a.ptr, nbytes, None))
    e1.record()
    rt_copy_ms = e0.elapsed_ms(e1)
    del big_a, big_b
    rng = np.random.default_rng(args.seed)
    svals = np.sort(rng.choice(np.arange(3, SMAX + 1), size=n, replace=False)).astype(np.uint64)
    for b in res:
        check(lib.gf2t_set_x(b.ptr, nw, None))
    P64 = C.POINTER(C.c_uint64)
    rr = np.full(n, R, dtype=np.uint64)
    kk = np.full(n, args.chain_steps, dtype=np.uint64)
    cres = (C.c_void_p * n)(*[b.ptr for b in res])
    ctmp = (C.c_void_p * n)(*[b.ptr for b in tmp])
    e0.record()
    check(lib.gf2t_square_fold_batch(cres, ctmp, rr.ctypes.data_as(P64), svals.ctypes.data_as(P64),
                                     kk.ctypes.data_as(P64), n, None))
    e1.record()
    ms = e0.elapsed_ms(e1)
    kernel, launches, per = chain_launches(args.chain_steps, args.sq_fuse != 0)
    launch_ms = ms / (args.chain_steps / per)
    gbs = 16 * nw * n / (launch_ms * 1e-3) / 1e9
    print(json.dumps({"mode": "squaring", "residues": n, "steps": args.chain_steps,
                      "sq_ppt": args.sq_ppt, "sq_order": args.sq_order, "sq_nt": args.sq_nt,
                      "sq_fuse": args.sq_fuse, "kernel": kernel, "launches": launches,
                      "squarings_per_launch": per, "chain_ms": ms,
                      "residue_steps_per_s": n * args.chain_steps / (ms * 1e-3),
                      "avg_launch_ms": launch_ms, "GB_per_s": gbs, "frac": gbs / HBM_PEAK_GBS,
                      "copy_bytes_moved": 2 * nbytes16, "copy_ms": copy_ms,
                      "copy_GB_per_s": 2 * nbytes16 / (copy_ms * 1e-3) / 1e9,
                      "runtime_copy_GB_per_s": 2 * nbytes / (rt_copy_ms * 1e-3) / 1e9}),
          flush=True)


def c4_pairs(seed: int):
    """64 operand pairs: sizes log-uniform integers in [2^16, 2^27] bits,
    i.i.d. uniform bits, and a fold s in [1, 68139920] per pair."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(64):
        ab = int(round(2 ** rng.uniform(16, 27)))
        bb = int(round(2 ** rng.uniform(16, 27)))
        s = int(rng.integers(1, SMAX + 1))
        a = rng.integers(0, 2**64, (ab + 63) // 64, dtype=np.uint64)
        b = rng.integers(0, 2**64, (bb + 63) // 64, dtype=np.uint64)
        if ab & 63:
            a[-1] &= np.uint64((1 << (ab & 63)) - 1)
        if bb & 63:
            b[-1] &= np.uint64((1 << (bb & 63)) - 1)
        out.append((a, ab, b, bb, s))
    return out


def mode_fft(args):
    """The Cantor-product batch alone (c4_measure), one JSON line."""
    from mersenne_trinomial_amd import device
    device.init(0)
    print(json.dumps({"mode": "fft", **c4_measure(args)}), flush=True)


def c4_measure(args):
    """The Cantor-product batch (64 seeded pairs of 2^16..2^27 bits, each
    product folded mod x^136279841 + x^s + 1): one pair at a time
    (gf2t_mulmod) and as one variable-size batch (gf2t_mulmod_batch), then
    2^24-point transforms with the forward and inverse timed apart, for
    the transform pass's rate, and full 136279841 x 136279841-bit modular
    products (wrapped: 2^23-point transforms plus a product of the top
    chunks)."""
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer, Event
    nw = (R + 63) // 64
    pairs = c4_pairs(args.seed)
    n = len(pairs)
    dev = []
    wsmax = 0
    for a, ab, b, bb, s in pairs:
        dev.append((DeviceBuffer.from_array(a), ab, DeviceBuffer.from_array(b), bb, s,
                    DeviceBuffer.words(nw)))
        wsmax = max(wsmax, int(lib.gf2t_mul_workspace_bytes(ab, bb)))
    work = DeviceBuffer(wsmax)
    P64 = C.POINTER(C.c_uint64)
    abits = np.array([p[1] for p in dev], dtype=np.uint64)
    bbits = np.array([p[3] for p in dev], dtype=np.uint64)
    svec = np.array([p[4] for p in dev], dtype=np.uint64)
    vp = C.c_void_p * n
    pa, pb, po = (vp(*[p[0].ptr for p in dev]), vp(*[p[2].ptr for p in dev]),
                  vp(*[p[5].ptr for p in dev]))
    bwork = DeviceBuffer(lib.gf2t_mulmod_batch_workspace_bytes(n, abits.ctypes.data_as(P64),
                                                               bbits.ctypes.data_as(P64)))

    def one_by_one():
        for da, ab, db, bb, s, do in dev:
            check(lib.gf2t_mulmod(da.ptr, ab, db.ptr, bb, R, s, do.ptr, work.ptr, None),
                  "gf2t_mulmod")

    def batch():
        check(lib.gf2t_mulmod_batch(n, pa, abits.ctypes.data_as(P64), pb,
                                    bbits.ctypes.data_as(P64), R, svec.ctypes.data_as(P64), po,
                                    bwork.ptr, None), "gf2t_mulmod_batch")

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        device.synchronize()
        e0, e1 = Event(), Event()
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        device.synchronize()
        return e0.elapsed_ms(e1) / args.steps

    ms_one = timed(one_by_one)
    ms_batch = timed(batch)
    bits = sum(ab + bb for _, ab, _, bb, _ in pairs)
    # full-size products: forward pair, pointwise, inverse timed with events
    plan = C.c_void_p()
    check(lib.gf2t_fft_plan_create(24, C.byref(plan)))
    fp, ip = C.c_int(), C.c_int()
    check(lib.gf2t_fft_plan_passes(plan, C.byref(fp), C.byref(ip)))
    f = DeviceBuffer.words(1 << 24)
    f.fill(0x3C)
    e = [Event() for _ in range(3)]
    reps = max(3, args.steps)
    check(lib.gf2t_fft_fwd(plan, f.ptr, None))
    check(lib.gf2t_fft_inv(plan, f.ptr, None))
    device.synchronize()
    e[0].record()
    for _ in range(reps):
        check(lib.gf2t_fft_fwd(plan, f.ptr, None))
    e[1].record()
    for _ in range(reps):
        check(lib.gf2t_fft_inv(plan, f.ptr, None))
    e[2].record()
    device.synchronize()
    fwd_ms, inv_ms = e[0].elapsed_ms(e[1]) / reps, e[1].elapsed_ms(e[2]) / reps
    lib.gf2t_fft_plan_destroy(plan)
    del f
    pass_bytes = 16 * (1 << 24)
    fwd_pass_ms = fwd_ms / fp.value
    mults = 24 * (1 << 23)            # GF(2^64) butterfly products per transform
    # full-size modular products (the deep scan's interval product)
    fa, fb, fo = pairs_full = (DeviceBuffer.words(nw), DeviceBuffer.words(nw),
                               DeviceBuffer.words(nw))
    rng = np.random.default_rng(args.seed)
    for x in (fa, fb):
        v = rng.integers(0, 2**64, nw, dtype=np.uint64)
        v[-1] &= np.uint64((1 << (R % 64)) - 1)
        x.upload(v)
    fwork = DeviceBuffer(lib.gf2t_mul_workspace_bytes(R, R) + 8 * (2 * nw + 8))

    def full():
        check(lib.gf2t_mulmod(fa.ptr, R, fb.ptr, R, R, 474, fo.ptr, fwork.ptr, None),
              "gf2t_mulmod")
    ms_full = timed(full)
    del pairs_full, fa, fb, fo, fwork
    return {"products": n, "operand_bits": bits,
            "one_by_one_ms": ms_one, "one_by_one_products_per_s": n / (ms_one * 1e-3),
            "batch_ms": ms_batch, "batch_products_per_s": n / (ms_batch * 1e-3),
            "batch_speedup": ms_one / ms_batch,
            "full_mulmod_ms": ms_full,
            "transform_2p24": {"fwd_ms": fwd_ms, "inv_ms": inv_ms,
                               "fwd_passes": fp.value, "inv_passes": ip.value,
                               "pass_ms": fwd_pass_ms,
                               "pass_GB_per_s": pass_bytes / (fwd_pass_ms * 1e-3) / 1e9,
                               "hbm_frac_per_pass": pass_bytes / (fwd_pass_ms * 1e-3)
                               / 1e9 / HBM_PEAK_GBS,
                               "gf64_mul_per_s": mults / (fwd_ms * 1e-3)}}


def mode_fft24(args):
    """Only 2^24-point transforms (forward then inverse, --steps times each)
    on one buffer, for rocprofv3 counter runs of k_fft_pass at the size of
    the full 136279841-bit product."""
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer
    device.init(0)
    plan = C.c_void_p()
    check(lib.gf2t_fft_plan_create(24, C.byref(plan)))
    f = DeviceBuffer.words(1 << 24)
    f.fill(0x3C)
    for _ in range(args.steps):
        check(lib.gf2t_fft_fwd(plan, f.ptr, None))
        check(lib.gf2t_fft_inv(plan, f.ptr, None))
    device.synchronize()
    fp, ip = C.c_int(), C.c_int()
    check(lib.gf2t_fft_plan_passes(plan, C.byref(fp), C.byref(ip)))
    lib.gf2t_fft_plan_destroy(plan)
    print(json.dumps({"mode": "fft24", "transforms": 2 * args.steps, "fwd_passes": fp.value,
                      "inv_passes": ip.value}), flush=True)


def mode_scan(args):
    """Deep scan of the given s values (one cohort) once per block size,
    with the scan's own phase times (device, gcd, resolution) on stderr."""
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    device.init(0)
    if args.scan_s.startswith("survivors:"):
        # the first N sieve survivors at depth k0 (s >= 3)
        n = int(args.scan_s.split(":")[1])
        smax = (args.scan_r - 1) // 2
        best, _ = M.run_gpu(args.scan_r, args.scan_k0, smax)
        sv = np.nonzero(best == np.uint64(0xFFFFFFFFFFFFFFFF))[0]
        svals = [int(x) for x in sv[sv >= 3][:n]]
    else:
        svals = [int(x) for x in args.scan_s.split(",") if x]
    for bm in [int(x) for x in args.scan_block.split(",") if x]:
        times = {}
        t0 = time.perf_counter()
        out = M.scan(args.scan_r, svals, k0=args.scan_k0, maxd=args.scan_maxd, L0=args.scan_l0,
                     Lmax=4096, verbose=1, block_m=bm, times=times)
        dt = time.perf_counter() - t0
        print(json.dumps({"mode": "scan", "r": args.scan_r, "k0": args.scan_k0,
                          "maxd": args.scan_maxd, "block_m": bm, "n_s": len(svals),
                          "graphs": os.environ.get("GF2T_SCAN_GRAPHS", "1") != "0",
                          "seconds": dt, "s_per_second": len(svals) / dt, **times,
                          "lines": [o.line() for o in out][:16]}), flush=True)


def mode_gcd(args):
    """Full-size interval gcds at r = 136279841: gcd(A, x^r + x^474 + 1) for
    random residues A (SplitMix64 words), one alone and then --gcd-conc at
    once on host threads (the deep scan's concurrency), with the half-gcd's
    device products.  Prints one JSON line with the wall times."""
    import threading
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import lib
    device.init(0)
    P64 = C.POINTER(C.c_uint64)
    nw = (R + 63) // 64
    t = np.zeros((R + 64) // 64, dtype=np.uint64)
    for e in (0, 474, R):
        t[e // 64] |= np.uint64(1 << (e % 64))

    def residue(seed):
        w = np.empty(nw, dtype=np.uint64)
        rng = np.random.default_rng(seed)     # PCG64: not GF(2)-linear
        w[:] = rng.integers(0, 2**64, nw, dtype=np.uint64)
        w[-1] &= np.uint64((1 << (R % 64)) - 1)
        return w

    def one(a, out, i):
        g = np.zeros(4, dtype=np.uint64)
        t0 = time.perf_counter()
        out[i] = (int(lib.gf2t_gcd_host(t.ctypes.data_as(P64), t.size, a.ctypes.data_as(P64),
                                        a.size, g.ctypes.data_as(P64), 4)),
                  time.perf_counter() - t0)

    res = {}
    a0 = residue(args.seed)
    out = [None]
    one(a0, out, 0)                               # warm plans and buffers
    one(a0, out, 0)
    res["one_s"] = out[0][1]
    res["one_degree"] = out[0][0]
    print(f"gcd: one alone {out[0][1]:.2f} s (degree {out[0][0]})", file=sys.stderr, flush=True)
    n = args.gcd_conc
    if n > 1:
        As = [residue(args.seed + 1 + i) for i in range(n)]
        outs = [None] * n
        ths = [threading.Thread(target=one, args=(As[i], outs, i)) for i in range(n)]
        t0 = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        wall = time.perf_counter() - t0
        res.update(concurrent=n, concurrent_wall_s=wall, concurrent_each_s=[o[1] for o in outs],
                   concurrent_per_gcd_s=wall / n)
        print(f"gcd: {n} at once {wall:.2f} s", file=sys.stderr, flush=True)
    print(json.dumps({"mode": "gcd", "r": R, "s": 474, **res}), flush=True)


def mode_sieve(args):
    """The coarse sieve alone: --warmup + --steps runs over every s <= (r-1)/2
    to --depth, each reported with its wall time and the device time of
    the top degrees (log table / classes+marking).  Depths 33..36 build the
    wide tables (about 205 GB of HBM at 34; 35 and 36 in 2 and 4 slices of that size)."""
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import SieveStats, check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer

    device.init(0)
    dbest = DeviceBuffer.words(SMAX + 1)
    stats = (SieveStats * 37)()
    walls = []
    for i in range(args.warmup + args.steps):
        dbest.fill(0xFF)
        device.synchronize()
        t0 = time.perf_counter()
        check(lib.gf2t_sieve(R, 2, args.depth, SMAX, dbest.ptr, stats), "gf2t_sieve")
        device.synchronize()
        dt = time.perf_counter() - t0
        top = [{"d": int(stats[d].d), "ms_table": round(stats[d].ms_table, 1),
                "ms_classes": round(stats[d].ms_classes + stats[d].ms_heavy, 1),
                "classes": int(stats[d].classes), "marking": int(stats[d].marking)}
               for d in range(max(2, args.depth - 3), args.depth + 1)]
        timed = i >= args.warmup
        if timed:
            walls.append(dt)
        print(f"sieve run {i} ({'timed' if timed else 'warmup'}) d=2..{args.depth} "
              f"wall {dt:.3f}s {top}", file=sys.stderr, flush=True)
    best = dbest.download(SMAX + 1)
    certified = int(np.count_nonzero(best[1:] != np.uint64(0xFFFFFFFFFFFFFFFF)))
    ms_deg = {int(stats[d].d): round(stats[d].ms_table + stats[d].ms_classes + stats[d].ms_heavy, 2)
              for d in range(2, args.depth + 1)}
    print(json.dumps({"mode": "sieve", "r": R, "smax": SMAX, "depth": args.depth,
                      "certified": certified, "survivors": SMAX - certified,
                      "wall_s": [round(w, 4) for w in walls],
                      "s_certified_per_s": certified / float(np.mean(walls)) if walls else None,
                      "ms_per_degree": ms_deg}), flush=True)


def mode_pipeline(args):
    """The whole search for one r: the coarse sieve over every s <= (r-1)/2
    to --depth, then the deep scan of its survivors (k0 = depth) to the
    Swan bound.  With r = 132049 and depth 20 this is the plain-method
    configuration (66,024 certificate lines).  Prints one JSON line."""
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import certs, device
    device.init(0)
    r = args.scan_r
    smax = (r - 1) // 2
    t0 = time.perf_counter()
    best, _ = M.run_gpu(r, args.depth, smax)
    t1 = time.perf_counter()
    lines = {int(ln.split()[0]): ln for ln in certs.sieve_lines(best)}
    surv = [int(x) for x in certs.survivors(best)]
    out = M.scan(r, surv, k0=args.depth, maxd=0, block_m=0, gcd_threads=16, verbose=1)
    t2 = time.perf_counter()
    kinds = {}
    for o in out:
        lines[o.s] = o.line()
        kinds[o.kind] = kinds.get(o.kind, 0) + 1
    text = "".join(lines[s] + "\n" for s in sorted(lines))
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    import hashlib
    print(json.dumps({"mode": "pipeline", "r": r, "depth": args.depth, "s_total": smax,
                      "lines": len(lines), "sieve_certified": len(lines) - len(surv),
                      "survivors": len(surv), "scan_found": kinds.get(1, 0),
                      "scan_u": kinds.get(2, 0), "scan_primitive": kinds.get(3, 0),
                      "scan_range": kinds.get(4, 0), "sieve_s": t1 - t0, "scan_s": t2 - t1,
                      "seconds": t2 - t0, "certified_per_s": len(lines) / (t2 - t0),
                      "md5": hashlib.md5(text.encode()).hexdigest()}), flush=True)


def serve_jobs(args):
    """The serving mix: r from {132049, 756839, 3021377, 136279841}, a seeded
    contiguous s slice of log-uniform length 10^3..10^6 (capped at
    --serve-max-len, and at 256 for r = 136279841), sieve bound 32, deep
    bound 2048, Poisson arrivals at --serve-rate."""
    from mersenne_trinomial_amd.serve import Job
    rng = np.random.default_rng(args.seed)
    t = 0.0
    jobs = []
    for i in range(args.serve_jobs):
        r = int(rng.choice([132049, 756839, 3021377, R]))
        hi_len = min(10 ** 6, args.serve_max_len, 256 if r == R else 10 ** 6)
        n = int(round(np.exp(rng.uniform(np.log(min(1000, hi_len)), np.log(hi_len)))))
        lo = int(rng.integers(1, (r - 1) // 2 - n + 2))
        t += float(rng.exponential(1.0 / args.serve_rate))
        jobs.append(Job(job_id=i, r=r, s_lo=lo, s_hi=lo + n - 1, depth=32, maxd=2048, arrival=t))
    return jobs


def mode_serve(args):
    """Poisson stream of mixed jobs through the request loop: p50/p99 job
    latency and s-values/sec; two jobs are re-run alone afterwards and
    their certificates compared with their lines in the mix."""
    import copy
    from mersenne_trinomial_amd.serve import RequestLoop
    jobs = serve_jobs(args)
    pristine = copy.deepcopy(jobs)
    iv = args.serve_interval
    loop = RequestLoop(block_m=0, L0=iv, Lmax=iv)
    for sv_r in sorted({j.r for j in jobs}):
        loop.server(sv_r)                     # slabs allocated before the clock starts
    t0 = time.perf_counter()
    pending = list(jobs)
    completed = []
    last_note = time.perf_counter()
    while pending or loop.busy():
        if time.perf_counter() - last_note > 10:
            last_note = time.perf_counter()
            state = {r: (sv.last_live, sv.last_queued) for r, sv in loop.servers.items()}
            print(f"t={last_note - t0:.0f}s: {len(completed)} jobs done, "
                  f"server (live, queued) by r: {state}", file=sys.stderr, flush=True)
        now = time.perf_counter() - t0
        while pending and pending[0].arrival <= now:
            j = pending.pop(0)
            j.arrival = t0 + j.arrival
            loop.submit(j)
        done = loop.pump(1)
        completed.extend(done)
        for j in done:
            print(f"job {j.job_id} r={j.r} n={j.s_hi - j.s_lo + 1} done in "
                  f"{j.done_at - j.arrival:.2f}s", file=sys.stderr, flush=True)
        if not loop.busy() and pending:
            time.sleep(max(0.0, min(0.05, pending[0].arrival - (time.perf_counter() - t0))))
    makespan = time.perf_counter() - t0
    for j in jobs:
        if not j.done_at:
            j.done_at = j.arrival
    lat = np.array([j.done_at - j.arrival for j in jobs])
    n_s = sum(j.s_hi - j.s_lo + 1 for j in jobs)
    loop.close()
    # the same jobs alone, one at a time
    checks = []
    print(f"all {len(jobs)} jobs done in {makespan:.1f}s; solo re-runs", file=sys.stderr,
          flush=True)
    for j0 in pristine[:2]:
        solo = RequestLoop(block_m=0, L0=iv, Lmax=iv)
        j0.arrival = time.perf_counter()
        solo.submit(j0)
        while solo.busy():
            solo.pump(1)
        checks.append(j0.certificate() == jobs[j0.job_id].certificate())
        solo.close()
    print(json.dumps({"mode": "serve", "jobs": len(jobs), "rate_per_s": args.serve_rate,
                      "s_values": n_s, "makespan_s": makespan, "s_per_second": n_s / makespan,
                      "latency_p50_s": float(np.percentile(lat, 50)),
                      "latency_p99_s": float(np.percentile(lat, 99)),
                      "mix": [[j.r, j.s_hi - j.s_lo + 1, round(j.done_at - j.arrival, 3)]
                              for j in jobs],
                      "solo_equals_mix": checks}), flush=True)


def apply_tuning(args):
    """Command-line knobs, applied after gf2t_init has loaded the tuning
    file (which would otherwise override the knobs it lists)."""
    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import check, lib
    device.init(0)
    if args.sq_ppt:
        check(lib.gf2t_tune(b"sq_ppt", args.sq_ppt))
    if args.sq_order >= 0:
        check(lib.gf2t_tune(b"sq_order", args.sq_order))
    if args.sq_nt >= 0:
        check(lib.gf2t_tune(b"sq_nt", args.sq_nt))
    if args.sq_fuse >= 0:
        check(lib.gf2t_tune(b"sq_fuse", args.sq_fuse))
    if args.scan_stagger is not None:
        check(lib.gf2t_tune(b"scan_stagger", args.scan_stagger))
    for kv in args.tune:
        key, _, val = kv.partition("=")
        check(lib.gf2t_tune(key.encode(), int(val)))


def chain_launches(steps: int, fuse: bool):
    """(kernel, launches) of a chain of `steps` squarings: with fusion,
    steps // 2 two-squaring launches (plus one single launch for an odd
    count, which the roofline line does not cover)."""
    if fuse and steps >= 2:
        return "k_square2_fold_batch", steps // 2, 2
    return "k_square_fold_batch", steps, 1


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(the deep scan shards by independent s values, the sieve by class ranges "
              "with one elementwise-min exchange of best[])", file=sys.stderr)
        sys.exit(2)
    apply_tuning(args)
    default_tuning = not (args.sq_ppt or args.sq_order >= 0 or args.sq_nt >= 0
                          or args.sq_fuse >= 0)
    if args.mode == "squaring":
        return mode_squaring(args)
    if args.mode == "fft":
        return mode_fft(args)
    if args.mode == "scan":
        return mode_scan(args)
    if args.mode == "pipeline":
        return mode_pipeline(args)
    if args.mode == "serve":
        return mode_serve(args)
    if args.mode == "sieve":
        return mode_sieve(args)
    if args.mode == "gcd":
        return mode_gcd(args)
    if args.mode == "fft24":
        return mode_fft24(args)
    if args.mode == "c3":
        from mersenne_trinomial_amd import device
        device.init(0)
        print(json.dumps({"mode": "c3", **c3_measure(args, None)}), flush=True)
        return

    from mersenne_trinomial_amd import device
    from mersenne_trinomial_amd._capi import SieveStats, check, lib
    from mersenne_trinomial_amd.device import DeviceBuffer, Event

    device.init(0)
    nw = (R + 63) // 64
    n = args.residues
    dbest = DeviceBuffer.words(SMAX + 1)
    res = [DeviceBuffer.words(nw) for _ in range(n)]
    tmp = [DeviceBuffer.words(nw) for _ in range(n)]
    cres = (C.c_void_p * n)(*[b.ptr for b in res])
    ctmp = (C.c_void_p * n)(*[b.ptr for b in tmp])
    P64 = C.POINTER(C.c_uint64)
    rr = np.full(n, R, dtype=np.uint64)
    kk = np.full(n, args.chain_steps, dtype=np.uint64)
    stats = (SieveStats * 37)()
    svals = None

    def one_step(timed: bool):
        nonlocal svals
        e0, e1, e2 = Event(), Event(), Event()
        e0.record()
        dbest.fill(0xFF)
        check(lib.gf2t_sieve(R, 2, args.depth, SMAX, dbest.ptr, stats), "gf2t_sieve")
        e1.record()
        if svals is None:
            best = dbest.download(SMAX + 1)
            surv = np.nonzero(best == np.uint64(0xFFFFFFFFFFFFFFFF))[0]
            surv = surv[surv >= 3]
            rng = np.random.default_rng(args.seed)
            svals = np.sort(rng.choice(surv, size=n, replace=False)).astype(np.uint64)
        for b in res:
            check(lib.gf2t_set_x(b.ptr, nw, None))
        ech0 = Event()
        ech0.record()
        check(lib.gf2t_square_fold_batch(cres, ctmp, rr.ctypes.data_as(P64),
                                         svals.ctypes.data_as(P64), kk.ctypes.data_as(P64),
                                         n, None), "gf2t_square_fold_batch")
        e2.record()
        device.synchronize()
        return e0.elapsed_ms(e1), ech0.elapsed_ms(e2), e0.elapsed_ms(e2)

    for _ in range(args.warmup):
        one_step(False)
    device.synchronize()
    t_start = time.perf_counter()
    sieve_ms, chain_ms, step_ms = [], [], []
    for _ in range(args.steps):
        a, b, c = one_step(True)
        sieve_ms.append(a)
        chain_ms.append(b)
        step_ms.append(c)
    device.synchronize()
    wall = time.perf_counter() - t_start

    best = dbest.download(SMAX + 1)
    certified = int(np.count_nonzero(best[1:] != np.uint64(0xFFFFFFFFFFFFFFFF)))
    survivors = SMAX - certified
    per_step_ms = 1e3 * wall / args.steps

    # after the timed region: the library's 16-byte stream copy over the
    # chain's live bytes (read + write = 16 nw n), the rate a pure
    # streaming kernel reaches on this box
    big_a = DeviceBuffer.words(nw * n)
    big_b = DeviceBuffer.words(nw * n)
    nbytes16 = (8 * nw * n) & ~15
    check(lib.gf2t_stream_copy(big_b.ptr, big_a.ptr, nbytes16, None))
    ec0, ec1 = Event(), Event()
    ec0.record()
    check(lib.gf2t_stream_copy(big_b.ptr, big_a.ptr, nbytes16, None))
    ec1.record()
    device.synchronize()
    copy_gbs = 2 * nbytes16 / (ec0.elapsed_ms(ec1) * 1e-3) / 1e9
    del big_a, big_b, res, tmp
    sieve_rate = certified / (float(np.mean(sieve_ms)) * 1e-3)
    c3 = c3_measure(args, dbest) if args.c3 else None
    value = c3["s_per_s"] if c3 else sieve_rate
    c4 = None
    if args.c4:
        print("c4: 64 seeded Cantor products, folded", file=sys.stderr, flush=True)
        c4 = c4_measure(args)

    kernel, launches, per = chain_launches(args.chain_steps, args.sq_fuse != 0)
    launch_ms = float(np.mean(chain_ms)) / (args.chain_steps / per)
    bytes_per_launch = 16 * nw * n
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    deg = [{"d": stats[d].d, "ms": round(stats[d].ms_table + stats[d].ms_classes + stats[d].ms_heavy, 2),
            "ms_table": round(stats[d].ms_table, 2), "classes": int(stats[d].classes),
            "marking": int(stats[d].marking)} for d in range(2, args.depth + 1)]
    line = {
        "metric": "s-values certified/sec at r=136279841; squaring+fold GB/s vs HBM peak",
        "value": value,
        "unit": "s-values certified/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": (f"r={R} deep-scan workload: {args.c3_n} s drawn by SplitMix64 (seed "
                         f"{args.c3_seed}) from [3, {SMAX}]; coarse sieve d=2..{args.depth} over "
                         f"all s, then the batched deep scan of its survivors to degree "
                         f"{args.c3_maxd} with an interval gcd every {args.c3_interval} degrees "
                         f"and per-s early exit; value = s certified / wall seconds of that pass"
                         if c3 else
                         f"r={R}: coarse sieve d=2..{args.depth} over all s in 1..{SMAX}; value = "
                         f"s certified by the sieve / sieve time"),
            "steps": (f"each timed step: coarse sieve d=2..{args.depth} over all s in 1..{SMAX} + "
                      f"{args.chain_steps} batched squaring+fold steps on {n} sieve survivors "
                      f"(17 MB residues, {per} squarings per launch); the roofline is the "
                      f"squaring kernel of these steps"),
            "r": R, "smax": SMAX, "depth": args.depth, "residues": n,
            "chain_steps": args.chain_steps, "seed": args.seed,
        },
        "c3": c3,
        "c4": c4,
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "copy_peak_measured": copy_gbs,
            "frac_of_copy": achieved / copy_gbs,
            "traffic": SQ_TRAFFIC_256.get(kernel) if (n == 256 and default_tuning) else None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "squarings_per_launch": per,
            "launches": launches,
            "avg_launch_ms": launch_ms,
        },
        "sieve": {"ms": float(np.mean(sieve_ms)), "certified": certified,
                  "survivors": survivors, "s_certified_per_s": sieve_rate, "per_degree": deg},
        "squaring_fold": {"ms": float(np.mean(chain_ms)),
                          "residue_steps_per_s": n * args.chain_steps / (np.mean(chain_ms) * 1e-3),
                          "GB_per_s": achieved},
        "event_ms_per_step": float(np.mean(step_ms)),
        "device": device.device_name(),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(certified, int(svals[0]), c3)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
