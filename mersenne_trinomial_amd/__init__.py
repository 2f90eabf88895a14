"""MI355X-native GF(2) trinomial smallest-factor search.

Hot path (hand-written gfx950 HIP kernels in csrc/, C ABI in
include/gf2t.h, built into libgf2t.so):
  (a) Zech-log coarse sieve            -> engine.run_gpu
  (b) squaring + trinomial fold chains -> engine.TrinGPU, square_fold
  (c) Cantor additive-FFT products     -> engine.DevFFT, mul, mulmod
Certificate formats live in certs.py.
"""
from .engine import (CantorBasis, DevFFT, GF2TError, TrinGPU, U64_MAX, bits_to_words,  # noqa: F401
                     fft_fwd, mul, mul_out_words, mulmod, primitive_poly, run_gpu,
                     square_fold)

__version__ = "0.1.0"
from . import certs  # noqa: E402,F401
from .scan import ScanOut, scan, swan_bound  # noqa: E402,F401
