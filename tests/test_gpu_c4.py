"""The Cantor-product batch: 64 seeded operand pairs of 2^16..2^27 bits
(log-uniform, i.i.d. bits), each product folded mod x^136279841 + x^s + 1
for a seeded s -- the workload bench.py --mode fft times.  Products with
up to 2^22 bits in total are compared with the oracle's product; larger
ones through residues modulo two degree-61 irreducibles (the reference's
bigtest check, src/cantor_ref.cpp:159-224).  Every fold is compared with
the oracle's linear fold of the verified product."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
R136 = 136279841


def _clmul_mod(x: int, y: int, q: int, d: int) -> int:
    p = 0
    while y:
        if y & 1:
            p ^= x
        x <<= 1
        y >>= 1
    for b in range(p.bit_length() - 1, d - 1, -1):
        if (p >> b) & 1:
            p ^= q << (b - d)
    return p


def test_c4_batch_products_and_folds():
    import bench
    import mersenne_trinomial_amd as M
    from mersenne_trinomial_amd import device
    device.init(0)
    qs = [O.find_irreducible(61, seed) for seed in (0x1234567, 0xC0FFEE)]
    pairs = bench.c4_pairs(20250101)
    assert len(pairs) == 64
    sizes = sorted(ab + bb for _, ab, _, bb, _ in pairs)
    assert sizes[0] >= 2 * (1 << 16) - 2 and sizes[-1] <= 2 * (1 << 27)
    direct = 0
    for i, (a, ab, b, bb, s) in enumerate(pairs):
        print(f"pair {i}: {ab} x {bb} bits, s={s}", flush=True)
        p = M.mul(a, ab, b, bb)
        if ab + bb <= 1 << 22:
            assert np.array_equal(p, O.fft_mul(a, ab, b, bb)), i
            direct += 1
        else:
            for q in qs:
                want = _clmul_mod(O.residue(a, ab, q, 61), O.residue(b, bb, q, 61), q, 61)
                assert O.residue(p, ab + bb, q, 61) == want, (i, hex(q))
        got = M.mulmod(a, ab, b, bb, R136, s)
        assert np.array_equal(got, O.reduce_fold(p, ab + bb - 1, R136, s)), i
    assert direct >= 10
