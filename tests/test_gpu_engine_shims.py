"""The C++ engine shims of include/gf2t_engine.hpp, compiled against the
C ABI the way reference code would recompile against them: TsGPU
(src/tsfactor.cu:207-391) through one blocked step of the scan, each
member's output against the oracle's squarings and products."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mersenne_trinomial_amd")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tsgpu_exe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("shim") / "tsgpu_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "host", "tsgpu_check.cpp"), "-o", str(exe),
                           "-L" + PKG, "-lgf2t", "-Wl,-rpath," + PKG])
    return str(exe)


def frob(r, s, k):
    return O.square_fold_steps(O.x_residue(r), r, s, k)


def esym(ys, r, s):
    nw = O.words(r)
    e = [np.zeros(nw, np.uint64) for _ in range(len(ys) + 1)]
    e[0][0] = 1
    for y in ys:
        for j in range(len(e) - 1, 0, -1):
            e[j] = e[j] ^ O.mulmod(e[j - 1], r, y, r, r, s)
    return e[1:]


def block_poly(ys, r, s):
    c = np.zeros(O.words(r), np.uint64)
    c[0] = 1
    for y in ys:
        t = y.copy()
        t[0] ^= np.uint64(2)
        c = O.mulmod(c, r, t, r, r, s)
    return c


@pytest.mark.parametrize("r,s,m,k", [(4423, 505, 8, 9), (4423, 1, 5, 0), (300007, 1201, 4, 33)])
def test_tsgpu_block_step_vs_oracle(tsgpu_exe, tmp_path, r, s, m, k):
    print(f"TsGPU shim r={r} s={s} m={m} k={k}", flush=True)
    p = subprocess.run([tsgpu_exe, str(r), str(s), str(m), str(k), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and "tsgpu ok" in p.stdout, p.stderr
    nw = O.words(r)

    def load(name):
        return np.fromfile(tmp_path / name, dtype=np.uint64)

    y0 = [frob(r, s, k + t) for t in range(m)]
    y1 = [frob(r, s, k + m + t) for t in range(m)]
    h0, h1 = load("h0.bin").reshape(m, nw), load("h1.bin").reshape(m, nw)
    for j, w in enumerate(esym(y0, r, s)):
        assert np.array_equal(h0[j], w), ("init_h_at", j)
    for j, w in enumerate(esym(y1, r, s)):
        assert np.array_equal(h1[j], w), ("square_all_step", j)
    c0, c1 = block_poly(y0, r, s), block_poly(y1, r, s)
    assert np.array_equal(load("c0.bin"), c0)
    assert np.array_equal(load("c1.bin"), c1)
    assert np.array_equal(load("a1.bin"), c0)
    a2 = O.mulmod(c0, r, c1, r, r, s)
    assert np.array_equal(load("a2.bin"), a2)
    full = O.fft_mul(a2, r, c1, r)
    got = load("p.bin")
    assert np.array_equal(got[:len(full)], full) and not got[len(full):].any()
    assert np.array_equal(load("red.bin"), O.mulmod(a2, r, c1, r, r, s))
    assert np.array_equal(load("sq.bin"), O.square_fold(h1[0], r, s))
