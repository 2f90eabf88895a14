"""The autotuner's output file: every chosen knob is written as a
`key value` line that gf2t_tune accepts (no GPU: the values are given)."""
import ctypes as C


def test_tune_file_lines_are_accepted(tmp_path):
    from mersenne_trinomial_amd import autotune
    from mersenne_trinomial_amd._capi import lib
    path = str(tmp_path / "tune.txt")
    fft = {12: {"chosen": 10}, 24: {"chosen": 12}}
    sq = {"sq_fuse": {"chosen": 1}, "sq_nt": {"chosen": 1}}
    blocks = [{"k": 17, "chosen": 8}, {"k": 27, "chosen": 32}]
    gcd = {"key": "gcd_dev_bits", "chosen": 131072}
    base = {"key": "gcd_base_bits", "chosen": 3072}
    autotune.write(path, fft, sq, blocks, "test device", gcd, base)
    lines = [l.split() for l in open(path) if not l.startswith("#")]
    assert ["gcd_base_bits", "3072"] in lines and ["gcd_dev_bits", "131072"] in lines
    assert ["fft_tile_24", "12"] in lines and ["block_m_27", "32"] in lines
    saved = {}
    for key, _ in lines:
        v = C.c_int64()
        assert lib.gf2t_tune_get(key.encode(), C.byref(v)) == 0, key
        saved[key] = v.value
    try:
        for key, value in lines:
            assert lib.gf2t_tune(key.encode(), int(value)) == 0, key
            v = C.c_int64()
            assert lib.gf2t_tune_get(key.encode(), C.byref(v)) == 0 and v.value == int(value), key
    finally:
        for key, value in saved.items():
            lib.gf2t_tune(key.encode(), value)
