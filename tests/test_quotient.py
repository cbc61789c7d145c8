"""The FP64 triangular solves divide each row by its diagonal through the correctly rounded
reciprocal and two exact-remainder corrections (kernels.hpp, quotient64). Run on the host
through ml_quotient64, the sequence must give the IEEE division bit for bit wherever the
corrections are exact: signed zeros, numerators from 2^-969 up and quotients below 2^1020
(compared as bit patterns, so -0 and +0 differ); below that, for subnormal and near-underflow
numerators, and near overflow, it stays within one ulp of the division (the stated limit)."""
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


def within_one_ulp(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all((a == b) | (np.abs(a - b) <= np.spacing(np.abs(b)))))


def test_signed_zeros_are_exact():
    den = np.array([1.0, 3.0, 7.25, 1e-3, 2.0 ** 600, 1e6])
    for z in (0.0, -0.0):
        num = np.full(len(den), z)
        assert same_bits(quotient(num, den), num / den)
    assert same_bits(quotient([-0.0], [2.0]), [-0.0])


def test_small_numerators_exact_from_the_stated_bound_within_one_ulp_below():
    den = np.array([1.0, 3.0, 7.25, 1e-3, 2.0 ** 600, 1e6, 0.7])
    exact = [2.0 ** -969, -(2.0 ** -969), 2.0 ** -968, 1e-290, -3.3e-250]
    num = np.repeat(exact, len(den))
    d = np.tile(den, len(exact))
    keep = np.abs(num / d) < 2.0 ** 1020
    assert same_bits(quotient(num[keep], d[keep]), (num / d)[keep])
    below = [5e-324, -5e-324, 2.0 ** -1060, -(2.0 ** -1040), 2.0 ** -1022, 2.0 ** -970]
    num = np.repeat(below, len(den))
    d = np.tile(den, len(below))
    with np.errstate(all="ignore"):
        assert within_one_ulp(quotient(num, d), num / d)


def test_large_quotients_exact_below_the_stated_bound():
    num = np.array([1e300, -1.7e300, 3.0, 1e200, 2.0 ** 1000])
    den = np.array([1e-5, 0.5, 2.0 ** -1000, 1e-100, 0.75])
    keep = np.abs(num / den) < 2.0 ** 1020
    assert keep.all()
    assert same_bits(quotient(num, den), num / den)


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
