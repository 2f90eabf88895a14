"""The oracle (CPU restatement) against golden vectors, known answers,
closed forms and brute force.  No GPU."""
import hashlib

import numpy as np
import pytest

import oracle as O

R136 = 136279841


def md5_lines(lines):
    return hashlib.md5(("\n".join(lines) + "\n").encode()).hexdigest()


# ------------------------------------------------------------------ sieve
def test_sieve_golden_r132049_d20():
    # `sieve_ref run 132049 20 66024` of the reference: 60,233 lines
    best, _ = O.sieve(132049, 20, 66024)
    lines = O.cert_lines(best)
    assert len(lines) == 60233
    assert md5_lines(lines) == "737739bafe9d565781900c8f45828440"


@pytest.mark.parametrize("r,depth,smax", [(136279841, 11, 30000), (74207281, 10, 20000),
                                          (101, 9, 100)])
def test_sieve_vs_brute_force(r, depth, smax):
    # the reference's own selftest cases (src/sieve_ref.cpp:176-180)
    a, _ = O.sieve(r, depth, smax)
    b = O.brute_force(r, depth, smax)
    assert np.array_equal(a, b)


def test_sieve_class_counts_are_irreducible_counts():
    _, cls = O.sieve(3021377, 20, 100, nthreads=4)
    for d in range(2, 21):
        assert int(cls[d]) == O.count_irreducible(d)
    # Moebius closed form against small hand values and the reference log
    assert [O.count_irreducible(d) for d in range(2, 9)] == [1, 2, 3, 6, 9, 18, 30]
    assert O.count_irreducible(30) == 35790267        # src/coarse_sieve.log:44
    assert O.count_irreducible(36) == 1908866960      # src/coarse_sieve.log:173


def test_sieve_published_count_d20_r136279841():
    # src/README.md:86-90: 913,104 factored among s < 10^6 at depth 20
    best, _ = O.sieve(R136, 20, 999999, nthreads=8)
    assert int(np.count_nonzero(best[1:] != O.NONE)) == 913104


def test_known_answers():
    # README.md:89-93 and src/tsfactor_handoff.md:337-338
    best, _ = O.sieve(R136, 8, 40)
    assert O.cert_lines(best[:34])[-1] == "33 5 p29"
    assert O.verify_factor(R136, 33, O.poly_from_hex("29"))
    assert O.verify_factor(R136, 474, O.poly_from_hex("3991ef16129"))
    assert not O.verify_factor(R136, 475, O.poly_from_hex("3991ef16129"))
    assert not O.verify_factor(R136, 33, O.poly_from_hex("2b"))


def test_primitive_poly_small():
    # x^2+x+1, x^3+x+1, x^4+x+1, x^5+x^2+1
    assert [O.find_primitive_poly(d) for d in (2, 3, 4, 5)] == [7, 11, 19, 37]


# --------------------------------------------------------------- residues
def _naive_square_mod(a, r, s):
    nw = O.words(r)
    sq = O.mul_naive(a, r, a, r)
    return O.reduce_bitserial(sq, 2 * r - 1, r, s)[:nw]


def _rand_res(rng, r):
    a = rng.integers(0, 2**63, O.words(r), dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, O.words(r), dtype=np.uint64)
    if r & 63:
        a[-1] &= np.uint64((1 << (r & 63)) - 1)
    return a


SQ_CASES = [(r, s) for r in (4423, 4159, 4097, 4229, 9689)
            for s in (1, 2, 3, 63, 64, 65, 129, r // 3, (r - 1) // 2 - 1, (r - 1) // 2)]


@pytest.mark.parametrize("r,s", SQ_CASES[::3] + [(31, 3), (67, 33), (129, 64)])
def test_square_fold_vs_schoolbook(r, s):
    rng = np.random.default_rng(r * 1000 + s)
    a = _rand_res(rng, r)
    assert np.array_equal(O.square_fold(a, r, s), _naive_square_mod(a, r, s))


def test_square_fold_known_answers():
    # T = x^31 + x^3 + 1 is primitive: x^(2^31) = x (src/trinomial_stage2.cu:95-97)
    assert list(O.square_fold_steps(O.x_residue(31), 31, 3, 31)) == [2]
    # F_32 at r=136279841 s=474 from the reference stage-2 driver
    f = O.square_fold_steps(O.x_residue(R136), R136, 474, 32)
    assert hashlib.md5(f.tobytes()).hexdigest() == "ff5094d371880811cb34b354da3af4a7"


def test_square_fold_linear():
    rng = np.random.default_rng(5)
    r, s = 9689, 1234
    a, b = _rand_res(rng, r), _rand_res(rng, r)
    assert np.array_equal(O.square_fold(a ^ b, r, s), O.square_fold(a, r, s) ^ O.square_fold(b, r, s))


def test_least_degree_detected_by_frobenius_gcd():
    # x^(2^41) + x shares the degree-41 factor with T at s=474 and no
    # degree-d factor for d < 41 dividing into it below 41 (r=136279841
    # is too large for a naive gcd here, so check mod the factor)
    P = O.poly_from_hex("3991ef16129")
    assert O.verify_factor(R136, 474, P)


def test_reduce_fold_matches_bitserial():
    rng = np.random.default_rng(7)
    for r, s in [(4423, 1200), (4097, 1), (4097, 2048), (131, 65)]:
        for L in (r, r + 1, r + s - 1, 2 * r - 1):
            v = rng.integers(0, 2**63, O.words(L), dtype=np.uint64)
            if L & 63:
                v[-1] &= np.uint64((1 << (L & 63)) - 1)
            assert np.array_equal(O.reduce_fold(v, L, r, s), O.reduce_bitserial(v, L, r, s))


# --------------------------------------------------------------- GF(2^64)
def test_gf64_mul_vs_bitserial():
    rng = np.random.default_rng(11)
    for a, b in rng.integers(0, 2**64, (2000, 2), dtype=np.uint64):
        assert O.gf64_mul(int(a), int(b)) == O.gf64_mul_bitserial(int(a), int(b))


def test_cantor_basis_chain():
    b = [int(x) for x in O.cantor_basis()]
    assert b[0] == 1
    for i in range(1, 64):
        assert O.gf64_mul_bitserial(b[i], b[i]) ^ b[i] == b[i - 1]
        assert b[i] & 1 == 0         # the reference's root choice
    # b[1..2^k] spans GF(2^(2^k)): b[1] lies in GF(4)
    assert O.gf64_mul_bitserial(O.gf64_mul_bitserial(b[1], b[1]), b[1]) == 1 or b[1] != 0


@pytest.mark.parametrize("m", [1, 2, 3, 4, 6, 8])
def test_fft_is_evaluation_at_cantor_points(m):
    rng = np.random.default_rng(m)
    f = rng.integers(0, 2**64, 1 << m, dtype=np.uint64)
    F = O.fft_fwd(f, m)
    basis = O.cantor_basis()
    for i in range(1 << m):
        assert int(F[i]) == O.poly_eval64(f, O.cantor_point(basis, i))
    assert np.array_equal(O.fft_inv(F, m), f)


@pytest.mark.parametrize("abits,bbits", [(1, 1), (31, 33), (64, 64), (1000, 2500), (4097, 3333),
                                         (100000, 70001)])
def test_fft_mul_vs_schoolbook(abits, bbits):
    rng = np.random.default_rng(abits + bbits)
    a = rng.integers(0, 2**64, O.words(abits), dtype=np.uint64)
    b = rng.integers(0, 2**64, O.words(bbits), dtype=np.uint64)
    if abits & 63:
        a[-1] &= np.uint64((1 << (abits & 63)) - 1)
    if bbits & 63:
        b[-1] &= np.uint64((1 << (bbits & 63)) - 1)
    p = O.fft_mul(a, abits, b, bbits)
    assert np.array_equal(p, O.mul_comb(a, abits, b, bbits))
    if abits * bbits < 5e6:
        assert np.array_equal(p, O.mul_naive(a, abits, b, bbits))


def test_residue_homomorphism():
    rng = np.random.default_rng(3)
    q = O.find_irreducible(61, 0xDEADBEEF)
    a = rng.integers(0, 2**64, 300, dtype=np.uint64)
    b = rng.integers(0, 2**64, 200, dtype=np.uint64)
    p = O.fft_mul(a, 300 * 64, b, 200 * 64)
    ra, rb = O.residue(a, 300 * 64, q, 61), O.residue(b, 200 * 64, q, 61)
    # r(a*b) = r(a) r(b) mod q, computed with a 2-word product
    pa = O.mul_comb(np.array([ra], np.uint64), 64, np.array([rb], np.uint64), 64)
    assert O.residue(p, 500 * 64, q, 61) == O.residue(pa, 128, q, 61)


# -------------------------------------------------------------------- gcd
def test_gcd_planted_factor():
    rng = np.random.default_rng(9)
    for _ in range(10):
        g = rng.integers(0, 2**64, 3, dtype=np.uint64); g[0] |= np.uint64(1)
        u = rng.integers(0, 2**64, 20, dtype=np.uint64)
        v = rng.integers(0, 2**64, 25, dtype=np.uint64)
        a = O.mul_comb(g, 192, u, 20 * 64)
        b = O.mul_comb(g, 192, v, 25 * 64)
        d, G = O.gcd(a, b)
        assert d >= O.deg(g)
        # g | G
        dg, _ = O.gcd(G, g)
        assert dg == O.deg(g)


def test_gcd_tiny_brute_force():
    for a in range(1, 64):
        for b in range(1, 64):
            d, g = O.gcd(np.array([a], np.uint64), np.array([b], np.uint64))
            best = 1
            for c in range(1, 64):
                # c divides a and b?
                da = O.gcd(np.array([a], np.uint64), np.array([c], np.uint64))[0]
                db = O.gcd(np.array([b], np.uint64), np.array([c], np.uint64))[0]
                if da == O.deg(np.array([c], np.uint64)) == db and c.bit_length() > best.bit_length():
                    best = c
            assert int(g[0]) == best
