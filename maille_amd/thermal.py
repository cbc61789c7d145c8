"""Transient heat conduction on the device: ``ThermalModel.solve_step`` mirrors
``Model::solve_step`` (model.rs:620-706) for a thermal model on tetrahedra, the transient caller
of the device PCG: the left matrix C/dt + theta K is assembled and factored once per step size,
and every increment is one device PCG solve.

A ``ThermalModel`` holds the mesh, per-element conductivity k and rho*c, convection faces
(element, face, film, sink), heat sources (element, q) and nodal heat loads, and fixed
temperatures (``fixed`` / ``imposed`` per node), as ``Model::with_analysis(.., Thermal, ..)``,
``with_heat_sources`` and ``with_convection`` define them (model.rs:257-470)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import FP64, SolveInfo, _ai64, _check, _d, _f64, _Handle, _i, _U8, lib
from .plasticity import PlasticityInputError


@dataclass(frozen=True)
class Step:
    """Step::transient (model.rs:125-134) and with_theta (:152-159): time_period > 0 in
    ``increments`` equal steps, theta in [0.5, 1] (1: implicit Euler)."""

    time_period: float
    increments: int
    theta: float = 1.0

    def __post_init__(self):
        if not (np.isfinite(self.time_period) and self.time_period > 0):
            raise PlasticityInputError(f"time period must be finite and > 0: {self.time_period}")
        if int(self.increments) <= 0:
            raise PlasticityInputError("a Step needs at least one increment")
        if not (0.5 <= self.theta <= 1.0):
            raise PlasticityInputError(f"theta must lie in [0.5, 1]: {self.theta}")

    @property
    def time_step(self) -> float:
        return self.time_period / float(self.increments)

    def with_theta(self, theta: float) -> "Step":
        return Step(self.time_period, self.increments, theta)


@dataclass
class ThermalIncrement:
    """An Increment of a transient Step: time, temperatures and reactions per node, heat fluxes
    at the Gauss points (ne, ngp, 6: x, y, z, qx, qy, qz)."""

    time: float
    temperatures: np.ndarray
    reactions: np.ndarray
    heat_fluxes: np.ndarray


class ThermalModel(_Handle):
    """The thermal model on the device (one temperature per node); close() frees it."""

    _destroy = "ml_thermal_destroy"

    def __init__(self, nodes, conn, conductivity, rho_c, fixed, imposed, loads=None,
                 heat_sources=(), convection=(), precision: int = FP64):
        nodes = _f64(nodes)
        conn = _ai64(conn)
        ne, nen = conn.shape
        self.nodes, self.conn = nodes, conn
        self.ne, self.nen, self.n_nodes = ne, nen, len(nodes)
        k = _f64(np.broadcast_to(np.asarray(conductivity, dtype=np.float64), (ne,)))
        rc = _f64(np.broadcast_to(np.asarray(rho_c, dtype=np.float64), (ne,)))
        self.loads = np.zeros(len(nodes)) if loads is None else _f64(loads).copy()
        self.heat_sources = list(heat_sources)
        self.convection = list(convection)
        fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
        imposed = _f64(imposed)
        nc, ce, cf, film, sink = self._conv()
        h = C.c_void_p()
        si = SolveInfo()
        _check(lib().ml_thermal_create(len(nodes), _d(nodes), ne, nen, _i(conn), _d(k), _d(rc),
                                       nc, _i(ce), _i(cf), _d(film), _d(sink),
                                       fixed.ctypes.data_as(_U8), _d(imposed), precision,
                                       C.byref(h), C.byref(si)), si)
        self._register(h)

    def _conv(self):
        conv = self.convection
        if not conv:
            return 0, np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64), np.zeros(1), \
                np.zeros(1)
        return (len(conv), _ai64([c[0] for c in conv]), _ai64([c[1] for c in conv]),
                _f64([c[2] for c in conv]), _f64([c[3] for c in conv]))

    def assemble_loads(self):
        """Model::assemble's F (model.rs:1089-1127): nodal loads, heat sources, convection."""
        f = self.loads.copy()
        src = self.heat_sources
        se = _ai64([s[0] for s in src]) if src else np.zeros(1, dtype=np.int64)
        sq = _f64([s[1] for s in src]) if src else np.zeros(1)
        nc, ce, cf, film, sink = self._conv()
        _check(lib().ml_thermal_add_loads(self._h, len(src), _i(se), _d(sq), nc, _i(ce), _i(cf),
                                          _d(film), _d(sink), _d(f)))
        return f

    def solve_step(self, step: Step, initial, fluxes: bool = True):
        """solve_step (model.rs:620-706): the increments from t = 0 (the initial field with the BC
        temperatures at the fixed nodes), and the solve info."""
        initial = _f64(np.broadcast_to(np.asarray(initial, dtype=np.float64), (self.n_nodes,)))
        f = self.assemble_loads()
        n = int(step.increments)
        ngp = 1 if self.nen == 4 else 4
        t = np.zeros((n + 1, self.n_nodes))
        r = np.zeros((n + 1, self.n_nodes))
        q = np.zeros((n + 1, self.ne, ngp, 6)) if fluxes else None
        si = SolveInfo()
        _check(lib().ml_thermal_solve_step(self._h, float(step.time_period), n,
                                           float(step.theta), _d(f), _d(initial), _d(t), _d(r),
                                           _d(q) if fluxes else None, C.byref(si)), si)
        dt = step.time_step
        return ([ThermalIncrement(time=k * dt, temperatures=t[k], reactions=r[k],
                                  heat_fluxes=q[k] if fluxes else None) for k in range(n + 1)],
                si.as_dict())

    def solve(self, fluxes: bool = True):
        """Model::solve for the thermal model (model.rs:548-610): K T = F on the device.
        Returns (ThermalIncrement at time 0 with the steady temperatures, reactions and heat
        fluxes, solve info)."""
        f = self.assemble_loads()
        ngp = 1 if self.nen == 4 else 4
        t = np.zeros(self.n_nodes)
        r = np.zeros(self.n_nodes)
        q = np.zeros((self.ne, ngp, 6)) if fluxes else None
        si = SolveInfo()
        _check(lib().ml_thermal_solve_steady(self._h, _d(f), _d(t), _d(r),
                                             _d(q) if fluxes else None, C.byref(si)), si)
        return ThermalIncrement(time=0.0, temperatures=t, reactions=r, heat_fluxes=q), si.as_dict()
