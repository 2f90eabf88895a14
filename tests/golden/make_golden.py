"""Write the golden fixtures in this folder from the reference's own code.

The reference's CPU sieve (src/sieve_ref.cpp) and host Cantor engine
(src/gf2_cantor_engine.h, src/gf2_gcd.h) are compiled where they lie by
oracle/Makefile into oracle/_ref/ (``make -C oracle ref REF=<reference>/src``);
this script runs them on fixed inputs and stores inputs and outputs as
data.  Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402


def splitmix64(state):
    """SplitMix64 stream (src/tsfactor.cu:1993-1998 uses the same generator)."""
    m = (1 << 64) - 1
    while True:
        state = (state + 0x9E3779B97F4A7C15) & m
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        yield z ^ (z >> 31)


def rand_poly(gen, bits):
    w = np.array([next(gen) for _ in range(O.words(bits))], dtype=np.uint64)
    if bits & 63:
        w[-1] &= np.uint64((1 << (bits & 63)) - 1)
    return w


def hexw(w):
    return [f"{int(x):016x}" for x in w]


def sieve_fixtures():
    for r, depth, smax in [(4423, 16, 2211), (132049, 14, 3000), (136279841, 24, 2000)]:
        lines = O.ref_sieve_lines(r, depth, smax)
        name = f"sieve_ref_r{r}_d{depth}_s{smax}.cert"
        with open(os.path.join(HERE, name), "w") as f:
            f.write("".join(ln + "\n" for ln in lines))
        print(name, len(lines))


def cantor_fixtures():
    ref = O.ref()
    basis = np.zeros(64, dtype=np.uint64)
    assert ref.ref_cantor_basis(O._ptr(basis)) == 64
    gen = splitmix64(0x5EED)
    gf64 = []
    for _ in range(256):
        a, b = next(gen), next(gen)
        gf64.append([f"{a:016x}", f"{b:016x}", f"{ref.ref_gf64_mul(a, b):016x}"])
    transforms = []
    for m in (1, 2, 3, 4, 6, 8):
        f = np.array([next(gen) for _ in range(1 << m)], dtype=np.uint64)
        g = f.copy()
        ref.ref_fft_fwd(O._ptr(g), m)
        transforms.append({"m": m, "in": hexw(f), "fwd": hexw(g)})
    products = []
    for abits, bbits in [(1, 1), (31, 33), (64, 64), (65, 127), (200, 1000), (1000, 2500),
                         (4096, 4095), (12345, 6789)]:
        a, b = rand_poly(gen, abits), rand_poly(gen, bbits)
        out = np.zeros(O.mul_out_words(abits, bbits), dtype=np.uint64)
        ref.ref_fft_mul(O._ptr(a), abits, O._ptr(b), bbits, O._ptr(out))
        products.append({"abits": abits, "bbits": bbits, "a": hexw(a), "b": hexw(b),
                         "out": hexw(out)})
    # a larger product kept as a digest of the reference's output
    abits, bbits = 1 << 20, (1 << 20) - 17
    a, b = rand_poly(gen, abits), rand_poly(gen, bbits)
    out = np.zeros(O.mul_out_words(abits, bbits), dtype=np.uint64)
    ref.ref_fft_mul(O._ptr(a), abits, O._ptr(b), bbits, O._ptr(out))
    big = {"abits": abits, "bbits": bbits, "seed": "continues the 0x5EED stream",
           "a_md5": hashlib.md5(a.tobytes()).hexdigest(),
           "b_md5": hashlib.md5(b.tobytes()).hexdigest(),
           "out_md5": hashlib.md5(out.tobytes()).hexdigest()}
    gcds = []
    for i in range(8):
        gbits = 40 + 30 * i
        g = rand_poly(gen, gbits)
        g[(gbits - 1) >> 6] |= np.uint64(1 << ((gbits - 1) & 63))
        ua, ub = rand_poly(gen, 300 + 17 * i), rand_poly(gen, 200 + 9 * i)
        a = O.mul_comb(g, gbits, ua, 300 + 17 * i)
        b = O.mul_comb(g, gbits, ub, 200 + 9 * i)
        gw = np.zeros(len(a) + 1, dtype=np.uint64)
        d = ref.ref_gcd_naive(O._ptr(a), len(a), O._ptr(b), len(b), O._ptr(gw), len(gw))
        gcds.append({"a": hexw(a), "b": hexw(b), "deg": int(d), "g": hexw(gw[:O.words(d + 1)])})
    with open(os.path.join(HERE, "cantor_ref_vectors.json"), "w") as f:
        json.dump({"source": "src/gf2_cantor_engine.h, src/gf2_gcd.h via oracle/ref_shim.cpp",
                   "basis": hexw(basis), "gf64_mul": gf64, "fft_fwd": transforms,
                   "products": products, "product_1m": big, "gcd_naive": gcds}, f, indent=0)
    print("cantor_ref_vectors.json")


if __name__ == "__main__":
    O.build(ref=True)
    if not O.have_ref():
        sys.exit("oracle/_ref is not built (the reference sources are needed)")
    sieve_fixtures()
    cantor_fixtures()


# The C2 digest in known_answers.json ("sieve_r3021377_d32_s1510688") comes
# from a run too long for this script (136.7 min on one core):
#     oracle/_ref/sieve_ref run 3021377 32 1510688 > c2.cert
# then md5sum c2.cert, wc -l c2.cert and the count of lines per degree.
