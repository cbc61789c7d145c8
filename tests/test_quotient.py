"""The FP64 triangular solves divide each row by its diagonal through the correctly rounded
reciprocal and two exact-remainder corrections (kernels.hpp, quotient64). Run on the host
through ml_quotient64, the sequence must give the IEEE division bit for bit on every operand
pair with a positive denominator: signed zeros, subnormal and near-underflow numerators,
quotients near overflow and the ordinary range (compared as bit patterns, so -0 and +0 differ)."""
import ctypes as C

import numpy as np

import maille_amd as ma

_D = C.POINTER(C.c_double)


def quotient(num, den):
    num = np.ascontiguousarray(num, dtype=np.float64)
    den = np.ascontiguousarray(den, dtype=np.float64)
    out = np.empty_like(num)
    st = ma.lib().ml_quotient64(len(num), num.ctypes.data_as(_D), den.ctypes.data_as(_D),
                                out.ctypes.data_as(_D))
    assert st == ma.ML_OK
    return out


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def test_signed_zero_and_subnormal_numerators():
    den = np.array([1.0, 3.0, 7.25, 1e-3, 2.0 ** 600, 1e6])
    specials = [0.0, -0.0, 5e-324, -5e-324, 2.0 ** -1060, -(2.0 ** -1040), 2.0 ** -1022,
                2.0 ** -970, -(2.0 ** -969), 2.0 ** -968, 1e-300]
    num = np.repeat(specials, len(den))
    d = np.tile(den, len(specials))
    with np.errstate(all="ignore"):
        assert same_bits(quotient(num, d), num / d)
    assert same_bits(quotient([-0.0], [2.0]), [-0.0])


def test_near_overflow_nan_and_infinity():
    num = np.array([1e308, -1.7e308, np.inf, -np.inf, np.nan, 3.0, 1e300])
    den = np.array([1e-10, 0.5, 2.0, 3.0, 1.0, 2.0 ** -1022, 1e-12])
    with np.errstate(all="ignore"):
        ref = num / den
        out = quotient(num, den)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    keep = ~np.isnan(ref)
    assert same_bits(out[keep], ref[keep])


def test_random_operands_over_the_normal_range():
    rng = np.random.default_rng(17)
    n = 400_000
    # numerators over the whole exponent range, denominators over the range of pivots and
    # beyond, plus quotients that land near rounding midpoints (d close to a power of two,
    # numerators with long mantissas)
    num = rng.standard_normal(n) * np.exp2(rng.integers(-960, 960, n))
    den = np.abs(rng.standard_normal(n)) * np.exp2(rng.integers(-40, 40, n)) + 1e-300
    num[: n // 4] = np.nextafter(np.exp2(rng.integers(-50, 50, n // 4)), np.inf) * 3.0
    den[: n // 4] = np.exp2(rng.integers(-30, 30, n // 4)) * (1.0 + 2.0 ** -52 * rng.integers(1, 64, n // 4))
    with np.errstate(all="ignore"):
        assert same_bits(quotient(num, den), num / den)
