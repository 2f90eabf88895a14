"""Loaders for the fixtures in tests/golden (written by
tests/golden/make_golden.py from the reference's own code, or copied as
data from the reference's published logs)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SIEVE_FIXTURES = [(4423, 16, 2211), (132049, 14, 3000), (136279841, 24, 2000)]


def sieve_lines(r, depth, smax):
    with open(os.path.join(GOLDEN, f"sieve_ref_r{r}_d{depth}_s{smax}.cert")) as f:
        return f.read().splitlines()


def known():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


def cantor():
    with open(os.path.join(GOLDEN, "cantor_ref_vectors.json")) as f:
        return json.load(f)


def words(hexes):
    return np.array([int(h, 16) for h in hexes], dtype=np.uint64)


def splitmix64(state):
    m = (1 << 64) - 1
    while True:
        state = (state + 0x9E3779B97F4A7C15) & m
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        yield z ^ (z >> 31)


def _rand_poly(gen, bits):
    w = np.array([next(gen) for _ in range((bits + 63) >> 6)], dtype=np.uint64)
    if bits & 63:
        w[-1] &= np.uint64((1 << (bits & 63)) - 1)
    return w


def product_1m_inputs():
    """Replay the generator of make_golden.py up to the 2^20-bit product."""
    v = cantor()
    gen = splitmix64(0x5EED)
    for _ in range(2 * len(v["gf64_mul"])):
        next(gen)
    for t in v["fft_fwd"]:
        for _ in range(1 << t["m"]):
            next(gen)
    for p in v["products"]:
        _rand_poly(gen, p["abits"])
        _rand_poly(gen, p["bbits"])
    big = v["product_1m"]
    return _rand_poly(gen, big["abits"]), _rand_poly(gen, big["bbits"]), big
