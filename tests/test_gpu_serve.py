"""Serving: the continuous-batching deep-scan server and the request loop.
Every s gets the same line as in the oracle's per-degree scan and as in
a solo run, whatever else shares the device and whenever it arrived."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    device.init(0)
    return M


def oracle_lines(r, svals, k0, maxd):
    want = {}
    for s in svals:
        kind, d, hx, _ = O.scan_one(r, s, k0, maxd)
        if kind:
            want[s] = f"{s} {d} p{hx}" if kind == 1 else f"{s} u"
    return want


def test_server_staggered_admission_small_capacity(M):
    from mersenne_trinomial_amd.serve import Server
    r, k0, maxd = 4423, 5, 200
    want = oracle_lines(r, range(2, 351), k0, maxd)
    svals = sorted(want)
    # room for a handful of slots only: most s wait in the queue and enter
    # as others finish, at different degree frontiers
    probe = Server(r)
    per = probe.bytes_per_slot
    probe.close()
    sv = Server(r, mem_bytes=7 * per, block_m=4, L0=16, Lmax=64)
    assert sv.slots == 7
    got = {}
    thirds = [svals[0::3], svals[1::3], svals[2::3]]
    sv.submit(1, thirds[0], k0=k0, maxd=maxd)
    for step in range(100000):
        if step == 3:
            sv.submit(2, thirds[1], k0=k0, maxd=maxd)
        if step == 11:
            sv.submit(3, thirds[2], k0=k0, maxd=maxd)
        live, queued = sv.step(1)
        for job, line in sv.poll():
            got[int(line.split()[0])] = line
        if step > 11 and live == 0 and queued == 0:
            break
    assert got == want
    sv.close()


def test_server_equals_scan_r132049(M):
    from mersenne_trinomial_amd.serve import Server
    r = 132049
    best, _ = O.sieve(r, 20, 66024)
    surv = [int(s) for s in np.nonzero(best == O.NONE)[0] if s >= 1][:200]
    ref = {o.s: o.line() for o in M.scan(r, surv, k0=20, maxd=2048)}
    sv = Server(r)
    sv.submit(7, surv[:120], k0=20, maxd=2048)
    got = {}
    steps = 0
    while True:
        if steps == 5:
            sv.submit(8, surv[120:], k0=20, maxd=2048)
        live, queued = sv.step(2)
        steps += 1
        got.update({int(ln.split()[0]): ln for _, ln in sv.poll()})
        if steps > 5 and live == 0 and queued == 0:
            break
    assert got == ref
    sv.close()


def test_request_loop_mix_equals_solo(M):
    from mersenne_trinomial_amd.serve import Job, RequestLoop
    specs = [(4423, 2, 900), (132049, 1000, 1600), (4423, 700, 1500), (132049, 30000, 30400)]

    def run(jobs):
        loop = RequestLoop(mem_fraction=0.3)
        for j in jobs:
            loop.submit(j)
        done = {}
        while loop.busy():
            for j in loop.pump(2):
                done[j.job_id] = j
        for j in jobs:
            done[j.job_id] = j
        loop.close()
        return {k: v.certificate() for k, v in done.items()}

    def mk(i, spec):
        r, lo, hi = spec
        return Job(job_id=i, r=r, s_lo=lo, s_hi=hi, depth=20, maxd=600)

    mixed = run([mk(i, sp) for i, sp in enumerate(specs)])
    for i, sp in enumerate(specs):
        solo = run([mk(i, sp)])
        assert solo[i] == mixed[i], sp
        assert len(mixed[i]) == sp[2] - sp[1] + 1


def test_request_loop_serving_mix_shape(M):
    """The serving mix's shape: jobs at r = 756839, 3021377 and 136279841
    (short slices), sieve bound 32, deep bound 2048 with one interval gcd
    per s.  Every job's certificate is the same in the mix as alone, and
    every factor line passes the oracle's check (divides x^r + x^s + 1,
    irreducible, of the stated degree)."""
    from mersenne_trinomial_amd.serve import Job, RequestLoop
    specs = [(756839, 1000, 1300), (3021377, 500000, 500200), (136279841, 20000000, 20000039)]

    def run(jobs):
        loop = RequestLoop(mem_fraction=0.5, L0=2048, Lmax=2048)
        for j in jobs:
            loop.submit(j)
        done = {}
        while loop.busy():
            for j in loop.pump(2):
                done[j.job_id] = j
        for j in jobs:
            done[j.job_id] = j
        loop.close()
        return {k: v.certificate() for k, v in done.items()}

    def mk(i, spec):
        r, lo, hi = spec
        return Job(job_id=i, r=r, s_lo=lo, s_hi=hi, depth=32, maxd=2048)

    print("serving mix: all jobs together", flush=True)
    mixed = run([mk(i, sp) for i, sp in enumerate(specs)])
    for i, sp in enumerate(specs):
        r, lo, hi = sp
        print(f"serving mix: r={r} alone", flush=True)
        solo = run([mk(i, sp)])
        assert solo[i] == mixed[i], sp
        assert len(mixed[i]) == hi - lo + 1
        deep = 0
        for ln in mixed[i]:
            f = ln.split()
            assert int(f[0]) in range(lo, hi + 1)
            if f[1] == "u":
                continue
            d, P = int(f[1]), O.poly_from_hex(f[2][1:])
            assert O.deg(P) == d and O.verify_factor(r, int(f[0]), P), ln
            deep += d > 32
        print(f"r={r}: {len(mixed[i])} lines, {deep} from the deep scan", flush=True)
