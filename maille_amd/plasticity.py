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
    NonPositive / NegativeProperty of the plasticity parameters). ``kind`` names the variant;
    ``point`` is the offending amplitude point for InvalidAmplitude."""

    def __init__(self, message, kind="invalid_property", point=None):
        super().__init__(message)
        self.kind = kind
        self.point = point


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
            raise PlasticityInputError("a Step needs at least one increment", "no_increments")

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
            raise PlasticityInputError("amplitude point 0 is invalid", "invalid_amplitude", 0)
        eps = np.finfo(np.float64).eps
        for k, (t, v) in enumerate(pts):
            ok = np.isfinite(t) and np.isfinite(v)
            ok = ok and ((abs(t) <= eps and abs(v) <= eps) if k == 0 else t > pts[k - 1][0])
            if not ok:
                raise PlasticityInputError(f"amplitude point {k} is invalid",
                                           "invalid_amplitude", k)
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


@dataclass
class Increment:
    """results.rs Increment for a quasi-static Step: time, displacements and reactions over all
    DOFs, Gauss-point stresses (ne, ngp, 9: position, then sxx, syy, szz, tyz, txz, txy), the
    equivalent plastic strain p (ne, ngp) and the relative residuals of the increment's Newton
    iterations (empty for the mesh at rest)."""

    time: float
    displacements: np.ndarray
    reactions: np.ndarray
    stresses: np.ndarray
    equivalent_plastic_strains: np.ndarray
    residuals: np.ndarray


def solve_quasi_static(model, step: Step, amplitude: Amplitude, yield_stress, hardening,
                       f=None, stresses: bool = True):
    """Model::solve_quasi_static (model.rs:715-797) on the device for ``model`` (a
    ``maille_amd.Model``): per-element yield stress (``np.inf`` for an elastic element) and
    hardening, loads ``f`` (default: the model's nodal loads, pressures and edge loads) and the
    model's BC values scaled by ``amplitude``. Returns (increments, info); raises
    ``NotConverged`` / ``PlasticCollapse`` like the reference."""
    from . import (QuasiStaticInfo, _check, _d, _f64, _i, lib)
    ne, nen, ndof = model.ne, model.nen, model.ndof
    ngp = 1 if nen == 4 else 4
    sy = _f64(np.broadcast_to(np.asarray(yield_stress, dtype=np.float64), (ne,)))
    hd = _f64(np.broadcast_to(np.asarray(hardening, dtype=np.float64), (ne,)))
    if np.any(np.isnan(sy)) or np.any(sy <= 0):
        raise PlasticityInputError("yield stress must be > 0 (infinite: elastic)")
    if np.any(~np.isfinite(hd)) or np.any(hd < 0):
        raise PlasticityInputError("hardening must be finite and >= 0")
    f = model.loads() if f is None else _f64(f)
    n = int(step.increments)
    at = _f64([t for t, _ in amplitude.points])
    av = _f64([v for _, v in amplitude.points])
    u = np.zeros((n + 1, ndof))
    r = np.zeros((n + 1, ndof))
    s = np.zeros((n + 1, ne, ngp, 9)) if stresses else None
    p = np.zeros((n + 1, ne, ngp))
    res = np.zeros((n, NEWTON_ITERATIONS + 1))
    cnt = np.zeros(n, dtype=np.int64)
    info = QuasiStaticInfo()
    st = lib().ml_model_solve_quasi_static(
        model._h, _d(sy), _d(hd), _d(f), float(step.time_period), n, len(at), _d(at), _d(av),
        _d(u), _d(r), _d(s) if stresses else None, _d(p), _d(res), _i(cnt), C.byref(info))
    try:
        _check(st, info)
    except Exception as err:
        # every failure names the Step increment it ended on, as NotConverged and
        # PlasticCollapse do in the reference
        err.increment = info.failed_increment
        err.cuts = info.cuts
        raise
    dt = step.time_step
    out = []
    for k in range(n + 1):
        out.append(Increment(time=(k * dt), displacements=u[k], reactions=r[k],
                             stresses=s[k] if stresses else None,
                             equivalent_plastic_strains=p[k],
                             residuals=res[k - 1, :cnt[k - 1]].copy() if k else np.zeros(0)))
    return out, info.as_dict()


def solve_large_strain(model, step: Step, amplitude: Amplitude, neo_hookean=False, pressures=(),
                       f=None):
    """Model::solve_quasi_static with Kinematics::Large (model.rs:732-1022) on the device for
    ``model`` (a ``maille_amd.Model``): Saint-Venant-Kirchhoff elements, or Neo-Hookean where
    ``neo_hookean`` (bool or per element) says so; ``pressures`` [(element, face, value)] follow
    the deformed faces. ``f`` defaults to the problem's dead loads: nodal loads and edge loads
    (pressures are follower loads here, as model.rs:1108-1114 leaves them out of F). Returns
    (increments, info); stresses are Cauchy stresses at the current Gauss-point positions."""
    from . import (QuasiStaticInfo, _ai64, _check, _d, _f64, _i, _U8, lib)
    ne, nen, ndof = model.ne, model.nen, model.ndof
    ngp = 1 if nen == 4 else 4
    neo = np.ascontiguousarray(np.broadcast_to(np.asarray(neo_hookean, dtype=np.uint8), (ne,)))
    if f is None:
        f = _f64(model.problem.f).copy()
        el = getattr(model.problem, "edge_loads", None)
        if el:
            e = _ai64([q[0] for q in el])
            ed = _ai64([q[1] for q in el])
            q = _f64(np.array([q[2] for q in el], dtype=np.float64).reshape(-1, 3))
            _check(lib().ml_model_add_edge_loads(model._h, len(e), _i(e), _i(ed), _d(q), _d(f)))
    else:
        f = _f64(f)
    pr = list(pressures)
    pe = _ai64([q[0] for q in pr]) if pr else np.zeros(1, dtype=np.int64)
    pf = _ai64([q[1] for q in pr]) if pr else np.zeros(1, dtype=np.int64)
    pv = _f64([q[2] for q in pr]) if pr else np.zeros(1)
    n = int(step.increments)
    at = _f64([t for t, _ in amplitude.points])
    av = _f64([v for _, v in amplitude.points])
    u = np.zeros((n + 1, ndof))
    r = np.zeros((n + 1, ndof))
    s = np.zeros((n + 1, ne, ngp, 9))
    res = np.zeros((n, NEWTON_ITERATIONS + 1))
    cnt = np.zeros(n, dtype=np.int64)
    info = QuasiStaticInfo()
    st = lib().ml_model_solve_large_strain(
        model._h, neo.ctypes.data_as(_U8), _d(f), len(pr), _i(pe), _i(pf), _d(pv),
        float(step.time_period), n, len(at), _d(at), _d(av), _d(u), _d(r), _d(s), _d(res),
        _i(cnt), C.byref(info))
    try:
        _check(st, info)
    except Exception as err:
        err.increment = info.failed_increment
        err.cuts = info.cuts
        raise
    dt = step.time_step
    return ([Increment(time=k * dt, displacements=u[k], reactions=r[k], stresses=s[k],
                       equivalent_plastic_strains=np.zeros((ne, ngp)),
                       residuals=res[k - 1, :cnt[k - 1]].copy() if k else np.zeros(0))
             for k in range(n + 1)], info.as_dict())
