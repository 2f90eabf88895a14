"""The oracle against the golden fixtures made from the reference's own
code (tests/golden/make_golden.py) and its published figures.  No GPU."""
import hashlib

import numpy as np
import pytest

import oracle as O
from golden_data import SIEVE_FIXTURES, cantor, known, sieve_lines, words


@pytest.mark.parametrize("r,depth,smax", SIEVE_FIXTURES)
def test_sieve_vs_reference_fixture(r, depth, smax):
    best, _ = O.sieve(r, depth, smax)
    assert O.cert_lines(best) == sieve_lines(r, depth, smax)


def test_sieve_reference_digest_r132049():
    g = known()["digests"]["sieve_r132049_d20_s66024"]
    lines = O.cert_lines(O.sieve(132049, 20, 66024)[0])
    assert len(lines) == g["lines"]
    assert hashlib.md5(("\n".join(lines) + "\n").encode()).hexdigest() == g["md5"]


def test_known_factor_lines():
    k = known()["known_factors"]
    for ln in k["lines"]:
        s, d, hx = ln.split()
        f = O.poly_from_hex(hx[1:])
        assert O.verify_factor(k["r"], int(s), f)
        assert O.deg(f) == int(d)


def test_s_below_1e6_count_d20():
    k = known()["r136279841_s_below_1e6"]
    best, _ = O.sieve(136279841, 20, k["smax"], nthreads=8)
    assert int(np.count_nonzero(best[1:] != O.NONE)) == k["factored_at_depth"]["20"]


def test_class_counts_are_irreducible_counts():
    for d, n in known()["r136279841_full_range"]["classes"].items():
        assert O.count_irreducible(int(d)) == n


def test_cantor_basis_and_gf64():
    v = cantor()
    assert np.array_equal(O.cantor_basis(), words(v["basis"]))
    for a, b, p in v["gf64_mul"]:
        assert O.gf64_mul(int(a, 16), int(b, 16)) == int(p, 16)
        assert O.gf64_mul_bitserial(int(a, 16), int(b, 16)) == int(p, 16)


def test_fft_forward_vectors():
    for t in cantor()["fft_fwd"]:
        assert np.array_equal(O.fft_fwd(words(t["in"]), t["m"]), words(t["fwd"])), t["m"]


def test_product_vectors():
    for p in cantor()["products"]:
        a, b, want = words(p["a"]), words(p["b"]), words(p["out"])
        assert np.array_equal(O.fft_mul(a, p["abits"], b, p["bbits"]), want)
        assert np.array_equal(O.mul_comb(a, p["abits"], b, p["bbits"]), want)


def test_gcd_vectors():
    for g in cantor()["gcd_naive"]:
        d, gw = O.gcd(words(g["a"]), words(g["b"]))
        assert d == g["deg"]
        assert np.array_equal(gw[:len(g["g"])], words(g["g"]))


def test_product_1m_digest():
    a, b, big = __import__("golden_data").product_1m_inputs()
    assert hashlib.md5(a.tobytes()).hexdigest() == big["a_md5"]
    assert hashlib.md5(b.tobytes()).hexdigest() == big["b_md5"]
    out = O.fft_mul(a, big["abits"], b, big["bbits"])
    assert hashlib.md5(out.tobytes()).hexdigest() == big["out_md5"]


def test_scan_selftest_counts_r4423():
    k = known()["scan_selftest_r4423"]
    kinds = [O.scan_one(k["r"], s, k["k0"], k["maxd"])[0]
             for s in range(k["s_range"][0], k["s_range"][1] + 1)]
    assert sum(1 for x in kinds if x) == k["survivors"]
    assert kinds.count(1) == k["factors"] and kinds.count(2) == k["u"]


def test_scan_one_reproduces_sieve_ref_lines_r4423():
    """Per-value pin of the oracle's per-degree scan: for every s <= 2211 at
    r = 4423, scan_one over degrees 2..16 gives exactly the line the
    reference's sieve_ref printed for that s (degree and least mask), and
    'no factor' for every s the sieve left."""
    want = {}
    import golden_data
    with open(golden_data.GOLDEN + "/sieve_ref_r4423_d16_s2211.cert") as f:
        for ln in f:
            s, d, hx = ln.split()
            want[int(s)] = (int(d), hx[1:])
    assert len(want) > 1500
    for s in range(1, 2212):
        kind, d, hx, _ = O.scan_one(4423, s, 1, 16)
        if s in want:
            assert kind == 1 and (d, hx) == want[s], s
        else:
            assert kind == 2, s
