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
