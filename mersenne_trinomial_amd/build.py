"""Build the gfx950 engine library (libgf2t.so) in-tree with hipcc.

Every HIP source is compiled for gfx950 only, with packed-FP32 generation
disabled, and linked into ``mersenne_trinomial_amd/libgf2t.so``.  The C++
command-line drivers under ``cli/`` are linked against it.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgf2t.so")
BIN = os.path.join(PKG, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["capi.hip", "square.hip", "fft.hip", "sieve.hip", "ddf.hip"]
HOST_SOURCES = ["host_poly.cpp", "hgcd.cpp", "scan.cpp", "serve.cpp"]
# command-line drivers with the reference's argv forms and output files
CLIS = ["coarse_sieve", "trinomial_stage2", "trinomial_scan", "tsfactor"]
CXX = os.environ.get("CXX", "g++")
HOSTFLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-mpclmul", "-msse4.1", "-pthread",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIPFLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + p.stdout + p.stderr)
        raise RuntimeError(f"compile failed: {cmd}")
    return p


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        jobs.append([HIPCC, *COMMON, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        jobs.append([CXX, *HOSTFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    clis = []
    for name in CLIS:
        out = os.path.join(BIN, name)
        clis.append([CXX, "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
                     os.path.join(PKG, "cli", name + ".cpp"), "-o", out + ".tmp", "-L" + PKG,
                     "-lgf2t", "-Wl,-rpath,$ORIGIN/.."])
    with ThreadPoolExecutor(max_workers=len(clis)) as ex:
        list(ex.map(_run, clis))
    for name in CLIS:
        os.replace(os.path.join(BIN, name + ".tmp"), os.path.join(BIN, name))
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
