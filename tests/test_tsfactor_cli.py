"""The tsfactor command-line drop-in (argv of src/tsfactor.cu:2119-2325,
result lines of :2480-2509): block-aligned intervals from --skip, the
last block's overshoot past --maxd, -z cut where the linear interval
schedule first exceeds Q blocks, the Swan bound, and resume from the
results file -- each line checked against the oracle's per-degree scan."""
import os
import random
import struct
import subprocess
import zlib

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


def _ckpt_check(tmp_path):
    exe = tmp_path / "ckpt_check"
    if not exe.exists():
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                               "-I" + os.path.join(ROOT, "mersenne_trinomial_amd", "cli"), "-o",
                               str(exe), os.path.join(ROOT, "tests", "host", "ckpt_check.cpp")])
    return str(exe)


def test_checkpoint_file_round_trip(tmp_path):
    """The checkpoint file: save/load round trip of per-s (s, k_done) pairs,
    the standard CRC-32 check value, rejection of a flipped bit and of a
    truncated file, and the writer keeping resumed s that have not run yet
    until their lines are flushed."""
    p = subprocess.run([_ckpt_check(tmp_path), str(tmp_path / "x.ckpt")], capture_output=True,
                       text=True)
    assert p.returncode == 0 and "ckpt ok" in p.stdout, p.returncode


def test_checkpoint_version1_still_loads(tmp_path):
    """A version-1 file (one k_done for every s) written byte by byte here
    loads as one pair per s."""
    # fields: version, r, skip, m, maxd, zq, k_done, n -- then n s values
    body = b"GF2TSCK1" + struct.pack("<8Q", 1, 4423, 8, 8, 203, 0, 72, 3) + \
        struct.pack("<3Q", 11, 13, 17)
    (tmp_path / "v1.ckpt").write_bytes(body + struct.pack("<I", zlib.crc32(body)))
    p = subprocess.run([_ckpt_check(tmp_path), "load", str(tmp_path / "v1.ckpt")],
                       capture_output=True, text=True)
    assert p.stdout.split() == ["load", "r=4423", "skip=8", "m=8", "maxd=203", "zq=0", "11:72",
                                "13:72", "17:72"], p.stdout


def _tsfc_bytes(r, s, skip, m, k_iv, blk, q, has_a, seed=1):
    """The reference's checkpoint file (src/tsfactor.cu:735-787) for a
    64-bit little-endian build: struct CkptHdr (twelve 8-byte fields, three
    u32 CRCs, padding to 112 bytes; crc_hdr over the first 108 bytes with
    itself zero), the m*nw-word h-table, then A when has_A."""
    nw = (r + 63) // 64
    rng = random.Random(seed)
    h = struct.pack(f"<{m * nw}Q", *[rng.getrandbits(64) for _ in range(m * nw)])
    a = struct.pack(f"<{nw}Q", *[rng.getrandbits(64) for _ in range(nw)]) if has_a else b""
    magic = 0x31434653544
    fields = struct.pack("<9QdQQ", magic, 1, r, s, skip, m, k_iv, blk, q, 0.0, has_a, nw)
    crc_h = zlib.crc32(h)
    crc_a = zlib.crc32(a) if has_a else 0
    hdr = fields + struct.pack("<III", crc_h, crc_a, 0)
    hdr = fields + struct.pack("<III", crc_h, crc_a, zlib.crc32(hdr)) + b"\0" * 4
    assert len(hdr) == 112
    return hdr + h + a


def test_reference_tsfc_checkpoint_is_read(tmp_path):
    """The TSFC reader takes the header fields of a reference checkpoint,
    checks all three CRCs, and falls back to the older generations
    (.ckpt.1, .ckpt.2) when the newest is damaged."""
    exe = _ckpt_check(tmp_path)
    base = tmp_path / "ts.ckpt"
    base.write_bytes(_tsfc_bytes(R, 505, 8, 8, 73, 2, 5, 1))
    p = subprocess.run([exe, "tsfc", str(base)], capture_output=True, text=True)
    f = dict(kv.split("=") for kv in p.stdout.split()[2:])
    assert p.stdout.startswith("tsfc " + str(base)), p.stdout
    assert f == {"r": "4423", "s": "505", "skip": "8", "m": "8", "k_iv": "73", "blk": "2",
                 "q": "5", "has_A": "1", "nw": "70"}
    # damage the A section of the newest file: the previous generation is used
    good1 = _tsfc_bytes(R, 580, 8, 8, 57, 0, 4, 0, seed=2)
    (tmp_path / "ts.ckpt.1").write_bytes(good1)
    bad = bytearray(base.read_bytes())
    bad[-3] ^= 4
    base.write_bytes(bytes(bad))
    p = subprocess.run([exe, "tsfc", str(base)], capture_output=True, text=True)
    assert p.stdout.startswith("tsfc " + str(base) + ".1 ") and "s=580" in p.stdout, p.stdout
    # a damaged header, and no other generation: invalid
    hb = bytearray(good1)
    hb[50] ^= 1
    (tmp_path / "ts.ckpt.1").write_bytes(bytes(hb))
    p = subprocess.run([exe, "tsfc", str(base)], capture_output=True, text=True)
    assert p.stdout.strip() == "tsfc invalid"


def test_results_file_long_lines(tmp_path):
    """A results line longer than any fixed buffer (a 4000-digit factor) is
    one line: only the s at the start of each line counts as done, so a
    hex run inside it that starts with a digit is not taken for an s."""
    hexd = "3" + "0123456789abcdef" * 250
    res = tmp_path / "w.results.txt"
    res.write_text(f"474 16001 p{hexd}\n505 u\n")
    p = subprocess.run([_ckpt_check(tmp_path), "done", str(res)], capture_output=True, text=True)
    assert p.stdout.split() == ["done", "474", "505"], p.stdout[:200]
    # the command-line drivers report the same count before they start
    # scanning (each on its own copy of the file: with a GPU present the
    # run then goes on and appends the line of s = 3)
    (tmp_path / "surv.txt").write_text("3\n474\n505\n")
    (tmp_path / "w2.results.txt").write_text(res.read_text())
    p = run(R, tmp_path / "surv.txt", "--skip", 8, "--out", tmp_path / "w", check=False)
    assert "2 already completed" in p.stderr, p.stderr
    p = subprocess.run([os.path.join(BIN, "trinomial_scan"), str(R), str(tmp_path / "surv.txt"),
                        "8", "100", str(tmp_path / "w2")], capture_output=True, text=True,
                       timeout=120)
    assert "3 survivors (2 already done)" in p.stderr, p.stderr


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


@pytest.mark.gpu
def test_resume_from_reference_tsfc_checkpoint(tmp_path):
    """An in-flight reference run's checkpoint (<out>.ckpt, one s at the
    start of an interval) is taken up: that s continues from the interval
    start, the others from --skip, and every line equals the straight run's."""
    skip, m, maxd = 8, 8, 203
    sv = survivors(2, 400, skip)
    last = skip + m * -(-(maxd - skip) // m)
    want = oracle_lines(sv, skip, last)
    # the s whose factor lies deepest, resumed at the interval start 8 blocks in
    deep = max(sv, key=lambda s: int(want[s].split()[1]) if " p" in want[s] else 0)
    k_iv = skip + 1 + 8 * m
    (tmp_path / "c.ckpt").write_bytes(_tsfc_bytes(R, deep, skip, m, k_iv, 0, 9, 0))
    (tmp_path / "surv.txt").write_text("".join(f"{s}\n" for s in sv))
    p = run(R, tmp_path / "surv.txt", "--skip", skip, "--maxd", maxd, "--out", tmp_path / "c")
    assert f"resuming s={deep} from" in p.stderr and f"clear through degree {k_iv - 1}" in p.stderr
    assert "m=8" in p.stderr
    lines = (tmp_path / "c.results.txt").read_text().splitlines()
    assert {int(ln.split()[0]): ln for ln in lines} == want
