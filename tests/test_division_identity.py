"""The ODE right-hand side divides by N as div_rn (dust2_amd/csrc/models.hpp):
q = RN(a * RN(1/b)), then RN(q + (a - q*b) * RN(1/b)) with the residual in
one fused multiply-add.  Markstein's theorem says this is the correctly
rounded a / b; checked here on the host against IEEE division, with the C
library's fma and fmaf (exact by definition), for random operands over a
wide exponent range and for divisors whose significand is all ones (the
case where a reciprocal is hardest to round)."""
import ctypes
import ctypes.util

import numpy as np
import pytest

_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.fma.restype = ctypes.c_double
_libm.fma.argtypes = [ctypes.c_double] * 3
_libm.fmaf.restype = ctypes.c_float
_libm.fmaf.argtypes = [ctypes.c_float] * 3


def _operands(rng, n, dtype, mant_bits, exp_span):
    ib = np.int64 if dtype == np.float64 else np.int32
    ub = np.uint64 if dtype == np.float64 else np.uint32
    bias = 1023 if dtype == np.float64 else 127
    mant = rng.integers(0, 1 << mant_bits, size=n, dtype=np.int64).astype(ub)
    expo = (bias + rng.integers(-exp_span, exp_span + 1, size=n)).astype(ub)
    sign = rng.integers(0, 2, size=n).astype(ub)
    width = 64 if dtype == np.float64 else 32
    bits = (sign << ub(width - 1)) | (expo << ub(mant_bits)) | mant
    return bits.astype(ub).view(dtype), ib


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_div_rn_is_the_correctly_rounded_quotient(dtype):
    rng = np.random.default_rng(7)
    n = 60000
    mant_bits = 52 if dtype == np.float64 else 23
    a, _ = _operands(rng, n, dtype, mant_bits, 60)
    b, _ = _operands(rng, n, dtype, mant_bits, 60)
    ub = np.uint64 if dtype == np.float64 else np.uint32
    # every third divisor with an all-ones significand; every fifth N = 1000
    ones = ub((1 << mant_bits) - 1)
    bb = b.view(ub).copy()
    bb[::3] |= ones
    b = bb.view(dtype)
    b[::5] = dtype(1000.0)
    fma = _libm.fma if dtype == np.float64 else _libm.fmaf
    inv = (dtype(1) / b).astype(dtype)
    q = (a * inv).astype(dtype)
    got = np.empty(n, dtype=dtype)
    for i in range(n):
        r = fma(-float(q[i]), float(b[i]), float(a[i]))
        got[i] = fma(r, float(inv[i]), float(q[i]))
    want = (a / b).astype(dtype)
    assert np.array_equal(got.view(ub), want.view(ub))
    # and the plain product is not always the quotient: the correction matters
    assert np.any(q.view(ub) != want.view(ub))
