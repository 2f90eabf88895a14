"""(a) coarse sieve on the GPU against the oracle, the reference's golden
output and closed forms.  Bit-exact best[] images."""
import hashlib

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
R136 = 136279841


@pytest.fixture(scope="module")
def eng():
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    device.init(0)
    return M


def md5_lines(lines):
    return hashlib.md5(("\n".join(lines) + "\n").encode()).hexdigest()


def test_golden_r132049_d20(eng):
    best, st = eng.run_gpu(132049, 20, 66024)
    lines = O.cert_lines(best)
    assert len(lines) == 60233
    assert md5_lines(lines) == "737739bafe9d565781900c8f45828440"
    for s in st:
        assert s.classes == O.count_irreducible(s.d)


@pytest.mark.parametrize("r,depth,smax", [(136279841, 11, 30000), (74207281, 10, 20000),
                                          (101, 9, 100), (5, 3, 2)])
def test_vs_brute_force(eng, r, depth, smax):
    best, _ = eng.run_gpu(r, depth, smax)
    assert np.array_equal(best, O.brute_force(r, depth, smax))


def test_published_count_d20(eng):
    best, _ = eng.run_gpu(R136, 20, 999999)
    assert int(np.count_nonzero(best[1:] != O.NONE)) == 913104
    ob, _ = O.sieve(R136, 20, 999999, nthreads=8)
    assert np.array_equal(best, ob)


def test_c2_shape_vs_oracle_d24(eng):
    best, _ = eng.run_gpu(3021377, 24, 1510688)
    ob, _ = O.sieve(3021377, 24, 1510688, nthreads=8)
    assert np.array_equal(best, ob)


def test_incremental_degrees_equal_one_shot(eng):
    a, _ = eng.run_gpu(756839, 22, 378419)
    b, _ = eng.run_gpu(756839, 16, 378419)
    c, _ = eng.run_gpu(756839, 22, 378419, best=b, min_depth=17)
    assert np.array_equal(a, c)


def test_high_degrees_marks_verified(eng):
    # d <= 28 at r=3021377: every class count equals the irreducible count,
    # the d <= 24 part matches the oracle, and sampled higher-degree marks
    # are genuine irreducible factors
    best, st = eng.run_gpu(3021377, 28, 1510688)
    for s in st:
        assert s.classes == O.count_irreducible(s.d), s.d
    ob, _ = O.sieve(3021377, 24, 1510688, nthreads=8)
    low = (best >> np.uint64(48)) <= 24
    assert np.array_equal(best[low], ob[low])
    assert np.all(ob[~low] == O.NONE)
    hi_s = np.nonzero(~low & (best != O.NONE))[0]
    rng = np.random.default_rng(0)
    for s in rng.choice(hi_s, size=min(200, hi_s.size), replace=False):
        k = int(best[s])
        d, mask = k >> 48, k & 0xFFFFFFFFFFFF
        assert O.is_irreducible_small(mask, d)
        assert O.divides_trinomial_small(3021377, int(s), mask, d)


def test_headline_full_range_d32_matches_reference_log(eng):
    # r=136279841, every s <= 68139920, degrees 2..32: the reference's
    # published run (src/coarse_sieve.log:39-64) leaves 4,110,974 survivors
    # at d<=29, 3,976,134 at d<=30, 3,849,633 at d<=31 and 3,731,849 at
    # d<=32, with 64,028,946 + 134,840 + 126,501 + 117,784 certificates
    smax = R136 // 2
    best, st = eng.run_gpu(R136, 32, smax)
    deg = best[1:] >> np.uint64(48)
    marked = best[1:] != O.NONE
    for d, surv in [(29, 4110974), (30, 3976134), (31, 3849633), (32, 3731849)]:
        assert int(np.count_nonzero(~(marked & (deg <= d)))) == surv, d
    assert int(np.count_nonzero(marked)) == 64028946 + 134840 + 126501 + 117784
    cls = {s.d: s.classes for s in st}
    assert cls[30] == 35790267 and cls[31] == 69273666 and cls[32] == 134215680
    # the two published known answers
    assert int(best[33]) == (5 << 48) | 0x29
    assert int(best[474]) == O.NONE   # its least factor has degree 41


@pytest.mark.parametrize("r,depth,smax", [(132049, 26, 66024), (3021377, 27, 200000)])
def test_partitioned_log_table_degrees_vs_oracle(eng, r, depth, smax):
    # degrees >= 26 build the log table by radix partition; the marks must
    # equal the oracle's direct-table sieve
    best, st = eng.run_gpu(r, depth, smax)
    ob, _ = O.sieve(r, depth, smax, nthreads=8)
    assert np.array_equal(best, ob)
    assert all(s.classes == O.count_irreducible(s.d) for s in st)


def test_wide_degrees_33_34_full_range_match_reference_log(eng):
    # degrees 33 and 34 over every s <= 68139920 at r=136279841, continuing
    # from the d<=32 image: the reference's runs (src/coarse_sieve.log:
    # 67-101) leave 3,620,029 survivors at d<=33 and 3,515,197 at d<=34,
    # with 111,820 and 104,832 new certificates; classes 260,300,986 and
    # 505,286,415; no heavy progressions.  Sampled new marks are genuine
    # irreducible factors of their trinomials.
    smax = R136 // 2
    base, _ = eng.run_gpu(R136, 32, smax)
    best, st = eng.run_gpu(R136, 34, smax, best=base, min_depth=33)
    deg = best[1:] >> np.uint64(48)
    marked = best[1:] != O.NONE
    assert int(np.count_nonzero(~(marked & (deg <= 33)))) == 3620029
    assert int(np.count_nonzero(~marked)) == 3515197
    assert int(np.count_nonzero(marked & (deg == 33))) == 111820
    assert int(np.count_nonzero(marked & (deg == 34))) == 104832
    got = {s.d: (s.classes, s.heavy) for s in st}
    assert got == {33: (260300986, 0), 34: (505286415, 0)}
    rng = np.random.default_rng(7)
    for d in (33, 34):
        idx = np.nonzero(marked & (deg == d))[0] + 1
        for s in rng.choice(idx, size=40, replace=False):
            k = int(best[s])
            mask = k & 0xFFFFFFFFFFFF
            assert O.is_irreducible_small(mask, d)
            assert O.divides_trinomial_small(R136, int(s), mask, d)
    # the d<=32 part is untouched
    assert np.array_equal(best[base != O.NONE], base[base != O.NONE])


@pytest.mark.parametrize("slices", [1, 2, 4, 8])
def test_sliced_log_table_equals_dense(eng, slices):
    """The 64-bit path with the log table built and used in value slices
    (how degrees 35 and 36 run) gives the same marks as the dense 32-bit
    path: degrees 25..27 at r = 136279841 forced through it with 1..8
    slices."""
    from mersenne_trinomial_amd._capi import lib
    smax = 1 << 21
    want, _ = eng.run_gpu(R136, 27, smax)
    try:
        assert lib.gf2t_tune(b"sieve_wide_from", 25) == 0
        assert lib.gf2t_tune(b"sieve_slices", slices) == 0
        print(f"sliced sieve: {slices} slices", flush=True)
        got, st = eng.run_gpu(R136, 27, smax)
    finally:
        assert lib.gf2t_tune(b"sieve_wide_from", 33) == 0
        assert lib.gf2t_tune(b"sieve_slices", 1) == 0
    assert np.array_equal(got, want)
    assert all(s.classes == O.count_irreducible(s.d) for s in st)
    assert lib.gf2t_tune(b"sieve_slices", 3) != 0


def test_wide_degrees_35_36_full_range_match_reference_log(eng):
    """Degrees 35 and 36 over every s <= 68139920 at r=136279841 (the log
    table in 2 and 4 slices), continuing from the d<=34 image: the
    reference's low-memory runs leave 3,415,883 survivors at d<=35 (the
    line count of its out35.survivors.txt) and 3,322,198 at d<=36, with
    93,685 new degree-36 certificates, 1,908,866,960 degree-36 classes
    and 2 heavy progressions (src/coarse_sieve.log:103-175).  Sampled new
    marks are genuine irreducible factors of their trinomials."""
    smax = R136 // 2
    base, _ = eng.run_gpu(R136, 34, smax)
    print("degrees 35..36", flush=True)
    best, st = eng.run_gpu(R136, 36, smax, best=base, min_depth=35)
    deg = best[1:] >> np.uint64(48)
    marked = best[1:] != O.NONE
    assert int(np.count_nonzero(~(marked & (deg <= 35)))) == 3415883
    assert int(np.count_nonzero(~marked)) == 3322198
    assert int(np.count_nonzero(marked & (deg == 36))) == 93685
    got = {s.d: (s.classes, s.heavy) for s in st}
    assert got[36] == (1908866960, 2)
    assert got[35][0] == O.count_irreducible(35)
    print({s.d: (round(s.ms_table, 1), round(s.ms_classes, 1)) for s in st}, flush=True)
    rng = np.random.default_rng(9)
    for d in (35, 36):
        idx = np.nonzero(marked & (deg == d))[0] + 1
        for s in rng.choice(idx, size=40, replace=False):
            k = int(best[s])
            mask = k & 0xFFFFFFFFFFFF
            assert O.is_irreducible_small(mask, d)
            assert O.divides_trinomial_small(R136, int(s), mask, d)
    assert np.array_equal(best[base != O.NONE], base[base != O.NONE])
