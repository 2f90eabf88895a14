"""Deep scan (device chains + host gcds) against the per-degree oracle:
the reference's own end-to-end selftest shape (src/trinomial_scan.cu:
476-535: r=4423, k0=5, maxd=200, s=2..350, default and stress interval
policies), then sieve survivors of r=132049."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    device.init(0)
    return M


def oracle_lines(r, svals, k0, maxd):
    want = {}
    for s in svals:
        kind, d, hx, _ = O.scan_one(r, s, k0, maxd)
        if kind == 0:
            continue
        want[s] = f"{s} {d} p{hx}" if kind == 1 else f"{s} u"
    return want


@pytest.mark.parametrize("L0,Lmax", [(32, 4096), (8, 8), (64, 64)])
def test_scan_r4423_vs_oracle(eng, L0, Lmax):
    r, k0, maxd = 4423, 5, 200
    want = oracle_lines(r, range(2, 351), k0, maxd)
    # src/README.md:107-112: 104 survivors, 102 factors and 2 "u"
    assert len(want) == 104
    assert sum(v.endswith(" u") for v in want.values()) == 2
    svals = sorted(want)
    got = eng.scan(r, svals, k0=k0, maxd=maxd, L0=L0, Lmax=Lmax)
    lines = {o.s: o.line() for o in got}
    assert lines == want


@pytest.mark.parametrize("parts", [2, 3, 7])
def test_scan_staggered_parts_vs_oracle(eng, parts):
    """A cohort split into staggered parts (gf2t_tune "scan_stagger": each
    part on its own thread and stream, starting its device work once the
    previous part has its first interval's products) gives the oracle's
    lines, with uneven part sizes and parts that finish at different
    intervals."""
    import ctypes as C
    from mersenne_trinomial_amd._capi import lib
    r, k0, maxd = 4423, 5, 200
    want = oracle_lines(r, range(2, 351), k0, maxd)
    svals = sorted(want)
    old = C.c_int64()
    assert lib.gf2t_tune_get(b"scan_stagger", C.byref(old)) == 0
    try:
        assert lib.gf2t_tune(b"scan_stagger", parts) == 0
        got = eng.scan(r, svals, k0=k0, maxd=maxd, L0=16, Lmax=64)
    finally:
        assert lib.gf2t_tune(b"scan_stagger", old.value) == 0
    assert {o.s: o.line() for o in got} == want
    assert lib.gf2t_tune(b"scan_stagger", 65) != 0


def test_scan_primitive_and_swan(eng):
    # x^127 + x^1 + 1 is primitive (127 is a Mersenne exponent, T(x)
    # irreducible): the scan reaches the bound with no factor
    out = eng.scan(127, [1, 7, 15, 30, 63], k0=1)
    by = {o.s: o for o in out}
    assert by[1].line() == "1 primitive"
    for s, o in by.items():
        kind, d, hx, _ = O.scan_one(127, s, 1, eng.swan_bound(127, s)[0])
        if kind == 1:
            assert o.line() == f"{s} {d} p{hx}"
        elif kind == 2:
            assert o.line() == f"{s} primitive"


def test_scan_sieve_survivors_r132049(eng):
    # survivors of the depth-20 sieve; every certificate verified as a
    # divisor and irreducible, and its degree checked minimal with the
    # per-degree oracle for the low-degree hits
    r = 132049
    best, _ = eng.run_gpu(r, 20, r // 2)
    surv = np.nonzero(best == np.uint64(0xFFFFFFFFFFFFFFFF))[0]
    surv = surv[surv >= 1]
    assert surv.size == 5791
    pick = surv[:96]
    out = eng.scan(r, pick, k0=20)
    checked_min = 0
    for o in out:
        assert o.kind == 1, o.line()
        P = O.poly_from_hex(o.hex)
        assert O.deg(P) == o.d
        assert O.verify_factor(r, o.s, P), o.line()
        if o.d <= 40 and checked_min < 12:
            kind, d, hx, _ = O.scan_one(r, o.s, 20, o.d)
            assert (kind, d, hx) == (1, o.d, o.hex), o.line()
            checked_min += 1
    assert checked_min >= 5


def test_scan_full_size_r136279841_s474(eng):
    # the deep scan at the full residue size: interval (32, 64] of
    # x^136279841 + x^474 + 1 holds its least factor, degree 41
    # (src/tsfactor_handoff.md:337-338); the interval gcd runs as a
    # half-gcd with device products
    got = eng.scan(136279841, [474], k0=32, maxd=64, L0=64, Lmax=64)
    assert [o.line() for o in got] == ["474 41 p3991ef16129"]
