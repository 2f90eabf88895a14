"""The plain-method configuration end to end on the device: x^132049 +
x^s + 1 for every s in 1..66024, coarse sieve to degree 20, then the deep
scan of the 5,791 survivors to the Swan bound (r/3, r/2 for s = 2).
66,024 certificate lines, checked against the reference's sieve digest,
the oracle's verifier (divisibility and irreducibility), and the oracle's
per-degree scan on a sample."""
import hashlib

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
R = 132049


@pytest.fixture(scope="module")
def c1():
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import certs, device
    device.init(0)
    best, _ = M.run_gpu(R, 20, 66024)
    lines = {int(ln.split()[0]): ln for ln in certs.sieve_lines(best)}
    surv = [int(x) for x in certs.survivors(best)]
    for o in M.scan(R, surv, k0=20, maxd=0, gcd_threads=16):
        lines[o.s] = o.line()
    return best, surv, lines


def test_every_s_has_one_line(c1):
    best, surv, lines = c1
    assert sorted(lines) == list(range(1, 66025))
    assert len(surv) == 5791


def test_sieve_part_is_the_reference_digest(c1):
    best, surv, lines = c1
    text = "".join(lines[s] + "\n" for s in sorted(lines) if best[s] != O.NONE)
    assert hashlib.md5(text.encode()).hexdigest() == "737739bafe9d565781900c8f45828440"


def test_scan_factors_verified(c1):
    """Every factor line of the deep scan, whatever its degree, passes the
    check-ntl criterion: the factor divides x^r + x^s + 1 and, up to the
    oracle's degree limit of 20,000, is irreducible (Rabin test).  The
    checks run on host threads (the oracle releases the interpreter)."""
    from concurrent.futures import ThreadPoolExecutor
    best, surv, lines = c1
    found = [lines[s] for s in surv if len(lines[s].split()) == 3]
    assert len(found) >= 5700

    def check(ln):
        s, d, hx = ln.split()
        f = O.poly_from_hex(hx[1:])
        return O.deg(f) == int(d) and O.verify_factor(R, int(s), f)

    # largest first, so the slow Rabin tests start early
    order = sorted(found, key=lambda ln: -int(ln.split()[1]))
    print(f"verifying {len(found)} factor lines, degrees up to {order[0].split()[1]}", flush=True)
    with ThreadPoolExecutor(16) as ex:
        bad = [ln for ln, ok in zip(order, ex.map(check, order)) if not ok]
    assert not bad, bad[:3]
    others = [lines[s] for s in surv if len(lines[s].split()) != 3]
    assert all(ln.endswith(" primitive") for ln in others)


def test_least_degree_sample_vs_oracle_scan(c1):
    best, surv, lines = c1
    rng = np.random.default_rng(5)
    small = [s for s in surv if len(lines[s].split()) == 3 and int(lines[s].split()[1]) <= 40]
    for s in rng.choice(small, 6, replace=False):
        d = int(lines[int(s)].split()[1])
        kind, d2, hx, _ = O.scan_one(R, int(s), 20, d)
        assert (kind, d2) == (1, d) and lines[int(s)] == f"{s} {d} p{hx}"


def test_slice_r136279841_256_certified():
    # a 256-s slice at r=136279841: sieve lines to degree 32, deep scan of
    # the survivors to degree 1024 ("s u" beyond); every factor line passes
    # the verifier (divides x^r + x^s + 1, irreducible)
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import certs
    r, lo = 136279841, 20250000
    best, _ = M.run_gpu(r, 32, lo + 255)
    seg = best[lo:lo + 256]
    lines = {}
    surv = []
    for k, key in enumerate(seg):
        s = lo + k
        if key == O.NONE:
            surv.append(s)
        else:
            lines[s] = certs.key_line(s, int(key))
    assert 0 < len(surv) < 64
    for o in M.scan(r, surv, k0=32, maxd=1024, L0=1024, Lmax=1024, gcd_threads=16):
        lines[o.s] = o.line()
    assert sorted(lines) == list(range(lo, lo + 256))
    for s, ln in lines.items():
        parts = ln.split()
        if len(parts) == 3:
            f = O.poly_from_hex(parts[2][1:])
            assert O.deg(f) == int(parts[1])
            assert O.verify_factor(r, s, f), ln
        else:
            assert parts[1] == "u", ln
