"""The tsfactor command-line drop-in (argv of src/tsfactor.cu:2119-2325,
result lines of :2480-2509): block-aligned intervals from --skip, the
last block's overshoot past --maxd, -z cut where the linear interval
schedule first exceeds Q blocks, the Swan bound, and resume from the
results file -- each line checked against the oracle's per-degree scan."""
import os
import subprocess

import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mersenne_trinomial_amd", "bin")
R = 4423


def run(*args, check=True, timeout=300):
    p = subprocess.run([os.path.join(BIN, "tsfactor"), *map(str, args)], capture_output=True,
                       text=True, timeout=timeout)
    if check and p.returncode != 0:
        raise AssertionError(f"tsfactor {args} -> {p.returncode}\n{p.stdout}\n{p.stderr}")
    return p


def oracle_lines(svals, skip, last, bound_kind="u"):
    """Lines for a scan of degrees skip+1..last: the least factor, else
    bound_kind ('u' when the bound was --maxd or -z, 'primitive' at the
    Swan bound)."""
    want = {}
    for s in svals:
        kind, d, hx, _ = O.scan_one(R, s, skip, last)
        assert kind, f"s={s} is not a survivor at depth {skip}"
        want[s] = f"{s} {d} p{hx}" if kind == 1 else f"{s} {bound_kind}"
    return want


def survivors(lo, hi, depth):
    return [s for s in range(lo, hi) if O.scan_one(R, s, depth, depth + 1)[0]]


def test_usage_and_required_skip(tmp_path):
    p = run(check=False)
    assert p.returncode == 1 and "Usage: tsfactor" in p.stderr
    (tmp_path / "s.txt").write_text("5\n")
    p = run(R, tmp_path / "s.txt", check=False)
    assert p.returncode == 1 and "--skip is REQUIRED" in p.stderr
    p = run(4424, tmp_path / "s.txt", "--skip", 8, check=False)
    assert p.returncode == 1 and "odd" in p.stderr
    p = run(R, tmp_path / "s.txt", "--skip", 8, "-z", 3, "-Z", 4, check=False)
    assert p.returncode == 1 and "mutually exclusive" in p.stderr


@pytest.mark.gpu
def test_tsfactor_selftest():
    p = run("--selftest")
    assert "TSFACTOR SELFTEST: ALL PASS" in p.stdout


@pytest.mark.gpu
def test_last_block_past_maxd_and_resume(tmp_path):
    # --maxd 203 with blocks of 8 from --skip 8: the last interval ends at
    # degree 208, so s = 505 and 580 (least factor of degree 207) get their
    # factor, where a per-degree scan to 203 would say 'u'
    skip, m, maxd = 8, 8, 203
    last = skip + m * -(-(maxd - skip) // m)
    assert last == 208
    sv = [s for s in survivors(2, 1200, skip) if s < 700]
    want = oracle_lines(sv, skip, last)
    assert want[505].startswith("505 207 p") and want[580].startswith("580 207 p")
    assert O.scan_one(R, 505, skip, maxd)[0] == 2
    surv = tmp_path / "surv.txt"
    surv.write_text("# survivors at depth 8\n" + "".join(f"{s}\n" for s in sv))
    first = sv[:5]
    (tmp_path / "t.results.txt").write_text("".join(want[s] + "\n" for s in first))
    p = run(R, surv, "--skip", skip, "--maxd", maxd, "--m", m, "--out", tmp_path / "t",
            "--batch", 64)
    assert "5 already completed" in p.stderr
    lines = (tmp_path / "t.results.txt").read_text().splitlines()
    assert len(lines) == len(sv)
    assert {int(ln.split()[0]): ln for ln in lines} == want
    # stdout carries the new lines in file order
    assert p.stdout.splitlines() == [want[s] for s in sv[5:]]


@pytest.mark.gpu
def test_z_stops_where_the_linear_schedule_passes_q(tmp_path):
    # --q0 2 blocks of 8, linear growth: intervals of 2, 4, 6 blocks; -z 5
    # gives up before the third, after degree 8 + (2 + 4) * 8 = 56
    skip, m = 8, 8
    sv = survivors(2, 500, skip)
    want = oracle_lines(sv, skip, skip + 6 * m)
    assert any(ln.endswith(" u") for ln in want.values())
    assert any(" p" in ln and int(ln.split()[1]) > 40 for ln in want.values())
    surv = tmp_path / "surv.txt"
    surv.write_text("".join(f"{s}\n" for s in sv))
    run(R, surv, "--skip", skip, "--m", m, "--q0", 2, "-z", 5, "--out", tmp_path / "z")
    lines = (tmp_path / "z.results.txt").read_text().splitlines()
    assert {int(ln.split()[0]): ln for ln in lines} == want
    # -Z caps q but keeps scanning: same lines as no cap at all
    run(R, surv, "--skip", skip, "--m", m, "--q0", 2, "-Z", 3, "--maxd", 120,
        "--out", tmp_path / "zz")
    lines = (tmp_path / "zz.results.txt").read_text().splitlines()
    assert {int(ln.split()[0]): ln for ln in lines} == oracle_lines(sv, skip, 120)


@pytest.mark.gpu
def test_swan_bound_primitive(tmp_path):
    # r = 4423 = -1 mod 8: for these s the Swan bound is r/3 = 1474;
    # x^4423 + x^271 + 1 and x^4423 + x^369 + 1 are irreducible
    skip, m = 8, 8
    sv = [271, 369, 1165, 505]
    for s in sv:   # Swan's condition: r = +-1 mod 8, s even not dividing 2r, or r - s so
        assert R % 8 in (1, 7) and (2 * R) % (R - s if s % 2 else s) != 0
    last = skip + m * -(-(1474 - skip) // m)
    want = oracle_lines(sv, skip, last, bound_kind="primitive")
    assert want[271] == "271 primitive" and want[369] == "369 primitive"
    surv = tmp_path / "surv.txt"
    surv.write_text("".join(f"{s}\n" for s in sv))
    p = run(R, surv, "--skip", skip, "--m", m, "--out", tmp_path / "w")
    lines = (tmp_path / "w.results.txt").read_text().splitlines()
    assert {int(ln.split()[0]): ln for ln in lines} == want
    assert "PRIMITIVE TRINOMIAL CANDIDATE: s=271" in p.stderr


def test_checkpoint_file_round_trip(tmp_path):
    """The checkpoint file: save/load round trip, the standard CRC-32 check
    value, and rejection of a flipped bit and of a truncated file."""
    exe = tmp_path / "ckpt_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                           "-I" + os.path.join(ROOT, "mersenne_trinomial_amd", "cli"), "-o", str(exe),
                           os.path.join(ROOT, "tests", "host", "ckpt_check.cpp")])
    p = subprocess.run([str(exe), str(tmp_path / "x.ckpt")], capture_output=True, text=True)
    assert p.returncode == 0 and "ckpt ok" in p.stdout, p.returncode


@pytest.mark.gpu
@pytest.mark.parametrize("prog", ["tsfactor", "trinomial_scan"])
def test_interrupted_run_resumes_from_checkpoint(tmp_path, prog):
    """A run stopped after its first interval writes the lines it has and a
    checkpoint (exit status 3); the same command then resumes the
    unfinished s from the checkpoint and ends with the straight run's lines."""
    skip, m, maxd = 8, 8, 203
    sv = [s for s in survivors(2, 700, skip)]
    if prog == "tsfactor":
        last = skip + m * -(-(maxd - skip) // m)
        args = [R, tmp_path / "surv.txt", "--skip", skip, "--maxd", maxd, "--m", m, "--out",
                tmp_path / "c"]
        binp = os.path.join(BIN, "tsfactor")
    else:
        last = maxd
        args = [R, tmp_path / "surv.txt", skip, maxd, tmp_path / "c"]
        binp = os.path.join(BIN, "trinomial_scan")
    want = oracle_lines(sv, skip, last)
    (tmp_path / "surv.txt").write_text("".join(f"{s}\n" for s in sv))

    def go(*extra):
        return subprocess.run([binp, *map(str, args), *map(str, extra)], capture_output=True,
                              text=True, timeout=300)

    p = go("--stop-after-intervals", 1)
    assert p.returncode == 3, p.stderr
    ck = tmp_path / "c.gf2t.ckpt"
    assert ck.exists()
    first = (tmp_path / "c.results.txt").read_text().splitlines()
    assert 0 < len(first) < len(sv)
    for ln in first:
        assert want[int(ln.split()[0])] == ln
    p = go()
    assert p.returncode == 0, p.stderr
    assert "resuming" in p.stderr and "clear through degree" in p.stderr
    assert not ck.exists()
    lines = (tmp_path / "c.results.txt").read_text().splitlines()
    assert len(lines) == len(sv)
    assert {int(ln.split()[0]): ln for ln in lines} == want
