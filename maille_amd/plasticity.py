"""Quasi-static elasto-plastic solves on the device: ``Step``, ``Amplitude`` and
``solve_quasi_static``, mirroring ``Model::solve_quasi_static`` (model.rs:715-797) for small
strains on tetrahedra with von Mises plasticity and linear isotropic hardening
(plasticity.rs:34-112). Every Newton-Raphson iteration runs on the device: the element
responses (return mapping, consistent tangent, internal forces), the gather assembly of the
tangent into the model's pattern, IC(0) with the shift ladder and the PCG solve."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

NEWTON_TOLERANCE = 1e-8  # model.rs:1220
NEWTON_ITERATIONS = 25   # model.rs:1222
MAX_CUTS = 5             # model.rs:1224


class PlasticityInputError(ValueError):
    """MailleError's input refusals of the quasi-static path (NoIncrements, InvalidAmplitude,
    NonPositive / NegativeProperty of the plasticity parameters)."""


@dataclass(frozen=True)
class Step:
    """Step::quasi_static (model.rs:125-150): ``time_period`` > 0 in ``increments`` equal
    steps."""

    time_period: float
    increments: int

    def __post_init__(self):
        if not (np.isfinite(self.time_period) and self.time_period > 0):
            raise PlasticityInputError(f"time period must be finite and > 0: {self.time_period}")
        if int(self.increments) <= 0:
            raise PlasticityInputError("a Step needs at least one increment")

    @property
    def time_step(self) -> float:
        """model.rs:171-173."""
        return self.time_period / float(self.increments)


class Amplitude:
    """Amplitude (model.rs:177-247): piecewise linear between (t, lambda) points, constant after
    the last one; the first point is (0, 0), times strictly increase, values are finite."""

    def __init__(self, points):
        pts = [(float(t), float(v)) for t, v in points]
        if not pts:
            raise PlasticityInputError("amplitude point 0 is invalid")
        eps = np.finfo(np.float64).eps
        for k, (t, v) in enumerate(pts):
            ok = np.isfinite(t) and np.isfinite(v)
            ok = ok and ((abs(t) <= eps and abs(v) <= eps) if k == 0 else t > pts[k - 1][0])
            if not ok:
                raise PlasticityInputError(f"amplitude point {k} is invalid")
        self.points = pts

    @classmethod
    def ramp(cls, time_period: float) -> "Amplitude":
        if not (np.isfinite(time_period) and time_period > 0):
            raise PlasticityInputError("ramp duration must be finite and > 0")
        return cls([(0.0, 0.0), (time_period, 1.0)])

    def at(self, t: float) -> float:
        pts = self.points
        after = 0
        while after < len(pts) and pts[after][0] <= t:
            after += 1
        if 1 <= after < len(pts):
            (t0, v0), (t1, v1) = pts[after - 1], pts[after]
            return v0 + (v1 - v0) * (t - t0) / (t1 - t0)
        if after >= 1:
            return pts[after - 1][1]
        return pts[0][1]


def von_mises_of(stresses) -> np.ndarray:
    """Equivalent stress (Stress::von_mises, results.rs:24-32) of Gauss-point records (..., 9): position, then
    sxx, syy, szz, tyz, txz, txy."""
    s = np.asarray(stresses)[..., 3:]
    a, b, c = s[..., 0], s[..., 1], s[..., 2]
    shear = s[..., 5] * s[..., 5] + s[..., 3] * s[..., 3] + s[..., 4] * s[..., 4]
    return np.sqrt(np.maximum(a * a + b * b + c * c - a * b - b * c - c * a + 3.0 * shear, 0.0))
