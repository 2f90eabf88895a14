"""Command-line drivers (coarse_sieve, trinomial_stage2, trinomial_scan):
argv forms and output files of src/coarse_sieve.cu:1273-1570,
src/trinomial_stage2.cu:1-311 and src/trinomial_scan.cu:541-690, checked
against the reference's golden sieve digest and the oracle."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mersenne_trinomial_amd", "bin")


def run(name, *args, check=True, timeout=300):
    p = subprocess.run([os.path.join(BIN, name), *map(str, args)], capture_output=True,
                       text=True, timeout=timeout)
    if check and p.returncode != 0:
        raise AssertionError(f"{name} {args} -> {p.returncode}\n{p.stdout}\n{p.stderr}")
    return p


@pytest.mark.parametrize("name", ["coarse_sieve", "trinomial_stage2", "trinomial_scan"])
def test_usage_without_arguments(name):
    args = ["usage"] if name == "trinomial_stage2" else []
    p = run(name, *args, check=False)
    assert p.returncode == 1
    assert "usage:" in p.stderr


def words(path):
    return np.fromfile(path, dtype=np.uint64)


@pytest.mark.gpu
def test_coarse_sieve_golden_r132049(tmp_path):
    pre = tmp_path / "c"
    run("coarse_sieve", 132049, 20, pre, "--found")
    cert = (tmp_path / "c.found.cert").read_bytes()
    assert hashlib.md5(cert).hexdigest() == "737739bafe9d565781900c8f45828440"
    surv = [int(x) for x in (tmp_path / "c.survivors.txt").read_text().split()]
    assert len(surv) == 66024 - 60233
    best = np.concatenate([[O.NONE], words(tmp_path / "c.found.bin")]).astype(np.uint64)
    assert O.cert_lines(best) == cert.decode().splitlines()
    assert surv == [int(s) for s in np.nonzero(best == O.NONE)[0] if s >= 1]


@pytest.mark.gpu
def test_coarse_sieve_continuation_equals_one_shot(tmp_path):
    # legacy argv form <r> <depth> <smax> <prefix>; deeper pass from the
    # survivors of a shallower one
    run("coarse_sieve", 756839, 22, 378419, tmp_path / "full", "--found")
    run("coarse_sieve", 756839, 14, 378419, tmp_path / "lo", "--found")
    run("coarse_sieve", 756839, 22, tmp_path / "hi", "--smax", 378419, "--min-depth", 15,
        "--load", tmp_path / "lo.found.bin", "--found")
    assert (tmp_path / "full.found.bin").read_bytes() == (tmp_path / "hi.found.bin").read_bytes()
    assert (tmp_path / "full.found.cert").read_bytes() == (tmp_path / "hi.found.cert").read_bytes()
    run("coarse_sieve", 756839, 22, tmp_path / "sv", "--min-depth", 15,
        "--load-survivors", tmp_path / "lo.survivors.txt")
    assert (tmp_path / "full.survivors.txt").read_bytes() == (tmp_path / "sv.survivors.txt").read_bytes()


@pytest.mark.gpu
def test_coarse_sieve_selftest():
    p = run("coarse_sieve", "--selftest")
    assert "selftest: PASS" in p.stdout


@pytest.mark.gpu
def test_stage2_selftest():
    p = run("trinomial_stage2", "selftest")
    assert "STAGE2 SELFTEST: ALL PASS" in p.stdout


@pytest.mark.gpu
def test_stage2_run_vs_oracle(tmp_path):
    r, s, k0, k1 = 4423, 1200, 5, 40
    run("trinomial_stage2", "run", r, s, k0, k1, tmp_path / "o")
    nw = (r + 63) // 64
    x = np.zeros(nw, np.uint64)
    x[0] = 2
    acc = np.zeros(nw, np.uint64)
    acc[0] = 1
    f = x
    for k in range(1, k1 + 1):
        f = O.square_fold(f, r, s)
        if k >= k0:
            t = f.copy()
            t[0] ^= np.uint64(2)
            acc = O.mulmod(acc, r, t, r, r, s)
    assert np.array_equal(words(tmp_path / "o.fk"), f)
    assert np.array_equal(words(tmp_path / "o.acc"), acc)
    assert (tmp_path / "o.meta").read_text().split() == [str(v) for v in (r, s, k0, k1)]


@pytest.mark.gpu
def test_stage2_resume_equals_straight_run(tmp_path):
    r, s = 132049, 4321
    run("trinomial_stage2", "run", r, s, 2000, 3000, tmp_path / "a")
    st = tmp_path / "state"
    run("trinomial_stage2", "run", r, s, 2000, 2100, tmp_path / "b", st)
    run("trinomial_stage2", "run", r, s, 2000, 3000, tmp_path / "b", st)
    for ext in ("acc", "fk"):
        assert (tmp_path / f"a.{ext}").read_bytes() == (tmp_path / f"b.{ext}").read_bytes()


@pytest.mark.gpu
def test_stage2_bench_cross_check():
    p = run("trinomial_stage2", "bench", 132049, 777)
    assert "spread-square == FFT-square: PASS" in p.stdout


@pytest.mark.gpu
def test_scan_selftest():
    p = run("trinomial_scan", "--selftest")
    assert "SCAN SELFTEST: ALL PASS" in p.stdout


@pytest.mark.gpu
def test_scan_cli_vs_oracle_and_resume(tmp_path):
    r, k0, maxd = 4423, 5, 200
    want = {}
    for s in range(2, 200):
        kind, d, hx, _ = O.scan_one(r, s, k0, maxd)
        if kind:
            want[s] = f"{s} {d} p{hx}" if kind == 1 else f"{s} u"
    sv = tmp_path / "surv.txt"
    sv.write_text("".join(f"{s}\n" for s in sorted(want)))
    first = sorted(want)[:10]
    (tmp_path / "o.results.txt").write_text("".join(want[s] + "\n" for s in first))
    p = run("trinomial_scan", r, sv, k0, maxd, tmp_path / "o", "--batch", 50)
    assert "(10 already done)" in p.stderr
    lines = (tmp_path / "o.results.txt").read_text().splitlines()
    assert len(lines) == len(want)
    assert {int(ln.split()[0]): ln for ln in lines} == want
