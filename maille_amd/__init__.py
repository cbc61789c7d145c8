"""maille_amd — MI355X implementation of maille-core's 3D linear-elastic solve path.

A ctypes face over ``libmaille_hip.so`` (include/maille_hip.h) that mirrors the reference's
public items for this path (crates/maille-core of timothee-faget/maille v0.14.0):

    Graph.from_elements / reverse_cuthill_mckee   iterative.rs:38, solve.rs:200
    Csr.zeros / Csr.add                            iterative.rs:91, :121
    Pcg(csr) / Pcg.solve                           iterative.rs:201, :210
    solid_d / solid_stiffness                      element.rs:193, :223
    Model(...).solve()                             model.rs:548

Errors mirror MailleError (error.rs): ``Singular`` and ``IterativeNotConverged`` are raised as
the exceptions of the same names. There is no CPU fallback: without the built library or a GPU
the calls that compute raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import itertools
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmaille_hip.so")
_lib = None

ML_OK, ML_SINGULAR, ML_NOT_CONVERGED, ML_INVALID_ARGUMENT, ML_OUT_OF_PATTERN, ML_HIP_ERROR, \
    ML_INVALID_MESH, ML_DEVICE_TIMEOUT, ML_INVALID_FILE, ML_NEWTON_NOT_CONVERGED, \
    ML_PLASTIC_COLLAPSE, ML_DIRECT_REQUIRED = range(12)
FP64, MIXED, MIXED_IC0, MIXED_JACOBI = 0, 1, 2, 3
AUTO, DIRECT, ITERATIVE = 0, 1, 2
TOLERANCE = 1e-12          # iterative.rs:9
MAX_ITERATIONS = 10_000    # iterative.rs:11
ITERATIVE_FROM = 100_000   # solve.rs:185

EXPORTS = (
    "ml_last_error", "ml_device_count", "ml_solid_d", "ml_stiffness_batch",
    "ml_stiffness_batch_noalloc", "ml_quotient64",
    "ml_pattern_build", "ml_pattern_n", "ml_pattern_nnz", "ml_pattern_get",
    "ml_pattern_ntargets", "ml_pattern_graph", "ml_pattern_free", "ml_layout_verify",
    "ml_layout_sweeps",
    "ml_pcg_create", "ml_pcg_solve", "ml_pcg_solve_device", "ml_pcg_multiply",
    "ml_pcg_multiply_many",
    "ml_pcg_precondition", "ml_pcg_precondition_many", "ml_pcg_factor", "ml_pcg_destroy",
    "ml_pcg_trace",
    "ml_pcg_set_wait_limit", "ml_pcg_set_sweeps", "ml_pcg_sweeps", "ml_model_set_sweeps",
    "ml_pcg_set_sweep_weight", "ml_pcg_sweep_weight", "ml_model_set_sweep_weight",
    "ml_model_sweep_weight",
    "ml_model_sweeps", "ml_model_index_bytes", "ml_model_spmv_index", "ml_model_residency",
    "ml_pcg_set_sweeps_fused", "ml_pcg_sweeps_fused", "ml_model_set_sweeps_fused",
    "ml_pcg_set_reference_dots", "ml_model_set_reference_dots",
    "ml_model_sweeps_fused",
    "ml_model_create", "ml_model_assemble", "ml_model_set_young", "ml_model_factor",
    "ml_model_solve", "ml_model_solve_device", "ml_model_stresses", "ml_model_stresses_device",
    "ml_model_nodal_stresses",
    "ml_model_add_pressures", "ml_model_add_edge_loads",
    "ml_model_nfree", "ml_model_nnz", "ml_model_levels", "ml_model_csr",
    "ml_model_compliance_gradient", "ml_model_destroy", "ml_model_profile",
    "ml_model_solve_many", "ml_batch_create", "ml_batch_solve", "ml_batch_destroy",
    "ml_pcg_autotune", "ml_model_autotune", "ml_pcg_batch_lanes", "ml_model_batch_lanes",
    "ml_pcg_set_batch_lanes", "ml_model_set_batch_lanes",
    "ml_msh_read", "ml_msh_sizes", "ml_msh_geometry", "ml_msh_set_info", "ml_msh_set_items",
    "ml_msh_destroy", "ml_model_solve_quasi_static",
    "ml_thermal_create", "ml_thermal_add_loads", "ml_thermal_solve_step", "ml_thermal_destroy",
    "ml_model_solve_large_strain", "ml_thermal_solve_steady",
)


class MailleError(Exception):
    """Base of the errors the library reports (error.rs:8)."""


class Singular(MailleError):
    """MailleError::Singular (error.rs:224)."""


class IterativeNotConverged(MailleError):
    """MailleError::IterativeNotConverged { iterations, residual } (error.rs:226-229)."""

    def __init__(self, iterations, residual):
        super().__init__(f"PCG did not converge: {iterations} iterations, residual {residual:e}")
        self.iterations = iterations
        self.residual = residual


class InvalidMesh(MailleError):
    """Mesh::new refusals (mesh.rs:257-348): ``kind`` and the offending ``element``."""

    KINDS = {1: "degenerate tetrahedron", 2: "inverted tetrahedron", 3: "curved Tet10 edge",
             4: "non-finite coordinate", 5: "unknown node", 6: "bad material", 7: "empty mesh"}

    def __init__(self, kind, element):
        super().__init__(f"{self.KINDS.get(kind, kind)} (element/node {element})")
        self.kind = kind
        self.element = element


class NotConverged(MailleError):
    """MailleError::NotConverged { increment, time } (error.rs:158-163): Newton-Raphson failed
    on the smallest cut of an increment of a quasi-static Step."""

    def __init__(self, increment, time):
        super().__init__(f"Newton-Raphson did not converge at increment {increment} (t = {time})")
        self.increment = increment
        self.time = time


class PlasticCollapse(MailleError):
    """MailleError::PlasticCollapse { increment, time } (error.rs:164-168): the tangent stopped
    being positive definite while Gauss points yield (a load beyond the limit load)."""

    def __init__(self, increment, time):
        super().__init__(f"plastic collapse at increment {increment} (t = {time})")
        self.increment = increment
        self.time = time


class DirectRequired(MailleError):
    """SolverKind::Auto below ITERATIVE_FROM free DOFs (solve.rs:185-194) is the direct skyline
    solver, which is not part of this library (the caller takes its own Direct arm)."""


class DeviceError(MailleError):
    """A HIP failure or an expired device wait."""


class QuasiStaticInfo(C.Structure):
    _fields_ = [("newton_iterations", C.c_int64), ("cuts", C.c_int64),
                ("pcg_iterations", C.c_int64), ("failed_increment", C.c_int64),
                ("failed_time", C.c_double), ("device_retries", C.c_int32),
                ("reserved", C.c_int32), ("seconds", C.c_double * 4)]

    def as_dict(self):
        return {"newton_iterations": self.newton_iterations, "cuts": self.cuts,
                "pcg_iterations": self.pcg_iterations, "failed_increment": self.failed_increment,
                "failed_time": self.failed_time, "device_retries": self.device_retries,
                "seconds": list(self.seconds)}


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("residual", C.c_double),
                ("true_residual", C.c_double), ("shift_index", C.c_int32),
                ("refinements", C.c_int32), ("bad_element", C.c_int64),
                ("mesh_error", C.c_int32), ("escalated", C.c_int32),
                ("device_retries", C.c_int32), ("reserved", C.c_int32),
                ("seconds", C.c_double * 6)]

    def as_dict(self):
        return {"iterations": self.iterations, "residual": self.residual,
                "true_residual": self.true_residual, "shift_index": self.shift_index,
                "refinements": self.refinements, "escalated": self.escalated,
                "device_retries": self.device_retries,
                "seconds": list(self.seconds)}


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)
_S = C.POINTER(SolveInfo)
_i64 = C.c_int64
_int = C.c_int


def lib():
    """The loaded library; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} is missing: run `python -m maille_amd.build` first")
        L = C.CDLL(_LIB_PATH)
        sig = {
            "ml_last_error": (C.c_char_p, []),
            "ml_device_count": (_int, []),
            "ml_solid_d": (_int, [C.c_double, C.c_double, _D]),
            "ml_stiffness_batch": (_int, [_int, _i64, _D, _D, _D, _D]),
            "ml_stiffness_batch_noalloc": (_int, [_int, _i64, _D, _D, _D, _D]),
            "ml_quotient64": (_int, [_i64, _D, _D, _D]),
            "ml_pattern_build": (_int, [_i64, _i64, _I, _I, _PP]),
            "ml_pattern_n": (_i64, [_P]),
            "ml_pattern_nnz": (_i64, [_P]),
            "ml_pattern_get": (_int, [_P, _I, _I, _I]),
            "ml_pattern_ntargets": (_i64, [_P]),
            "ml_pattern_graph": (_int, [_P, _I, _I]),
            "ml_pattern_free": (None, [_P]),
            "ml_layout_verify": (_int, [_i64, _I, _I, _D, _D, _D, _I]),
            "ml_layout_sweeps": (_int, [_i64, _I, _I, _int, _D, _D, _int, _D, _I]),
            "ml_pcg_create": (_int, [_i64, _I, _I, _D, _int, _PP]),
            "ml_pcg_solve": (_int, [_P, _D, _D, _i64, C.c_double, _i64, _S]),
            "ml_pcg_solve_device": (_int, [_P, _P, _P, C.c_double, _i64, _S]),
            "ml_pcg_multiply": (_int, [_P, _D, _D]),
            "ml_pcg_multiply_many": (_int, [_P, _i64, _D, _D]),
            "ml_pcg_precondition": (_int, [_P, _D, _D]),
            "ml_pcg_precondition_many": (_int, [_P, _i64, _D, _D]),
            "ml_pcg_factor": (_int, [_P, _D, _I32]),
            "ml_pcg_destroy": (None, [_P]),
            "ml_pcg_trace": (_int, [_P, _D, _I, _I, _I, _i64]),
            "ml_pcg_set_wait_limit": (_int, [_P, _i64]),
            "ml_pcg_set_sweeps": (_int, [_P, _int]),
            "ml_pcg_sweeps": (_int, [_P]),
            "ml_model_set_sweeps": (_int, [_P, _int]),
            "ml_model_sweeps": (_int, [_P]),
            "ml_pcg_set_sweep_weight": (_int, [_P, C.c_double]),
            "ml_pcg_sweep_weight": (C.c_double, [_P]),
            "ml_model_set_sweep_weight": (_int, [_P, C.c_double]),
            "ml_model_sweep_weight": (C.c_double, [_P]),
            "ml_model_index_bytes": (_int, [_P, _I32]),
            "ml_pcg_set_reference_dots": (_int, [_P, _int]),
            "ml_model_set_reference_dots": (_int, [_P, _int]),
            "ml_model_spmv_index": (_int, [_P, _I]),
            "ml_model_residency": (_int, [_P, _I32]),
            "ml_pcg_set_sweeps_fused": (_int, [_P, _int]),
            "ml_pcg_sweeps_fused": (_int, [_P]),
            "ml_model_set_sweeps_fused": (_int, [_P, _int]),
            "ml_model_sweeps_fused": (_int, [_P]),
            "ml_model_create": (_int, [_i64, _D, _i64, _int, _I, _D, _D, _U8, _D, _int, _int,
                                       _PP, _S]),
            "ml_model_assemble": (_int, [_P]),
            "ml_model_set_young": (_int, [_P, _D]),
            "ml_model_factor": (_int, [_P, _S]),
            "ml_model_solve": (_int, [_P, _D, _D, _D, _S]),
            "ml_model_solve_device": (_int, [_P, _P, _P, _P, _S]),
            "ml_model_stresses": (_int, [_P, _D, _D]),
            "ml_model_stresses_device": (_int, [_P, _P, _P]),
            "ml_model_nodal_stresses": (_int, [_P, _D, _D]),
            "ml_model_add_pressures": (_int, [_P, _i64, _I, _I, _D, _D]),
            "ml_model_add_edge_loads": (_int, [_P, _i64, _I, _I, _D, _D]),
            "ml_model_nfree": (_i64, [_P]),
            "ml_model_nnz": (_i64, [_P]),
            "ml_model_levels": (_i64, [_P]),
            "ml_model_csr": (_int, [_P, _I, _I, _I, _D]),
            "ml_model_compliance_gradient": (_int, [_P, _D]),
            "ml_model_destroy": (None, [_P]),
            "ml_model_profile": (_int, [_P, _int, _D, _I]),
            "ml_model_solve_many": (_int, [_P, _i64, _D, _D, _S]),
            "ml_batch_create": (_int, [_i64, C.POINTER(C.c_void_p), _int, _PP]),
            "ml_batch_solve": (_int, [_P, C.POINTER(_D), C.POINTER(_D), _S]),
            "ml_batch_destroy": (None, [_P]),
            "ml_pcg_autotune": (_int, [_P, _i64, _I32]),
            "ml_model_autotune": (_int, [_P, _i64, _I32]),
            "ml_pcg_batch_lanes": (_int, [_P]),
            "ml_pcg_set_batch_lanes": (_int, [_P, _int]),
            "ml_model_set_batch_lanes": (_int, [_P, _int]),
            "ml_model_batch_lanes": (_int, [_P]),
            "ml_msh_read": (_int, [_P, _i64, _PP, _I32, _I]),
            "ml_msh_sizes": (_int, [_P, _I]),
            "ml_msh_geometry": (_int, [_P, _D, _I]),
            "ml_msh_set_info": (_int, [_P, _i64, C.c_char_p, _i64, _I]),
            "ml_msh_set_items": (_int, [_P, _i64, _I, _I, _I, _I]),
            "ml_msh_destroy": (None, [_P]),
            "ml_thermal_create": (_int, [_i64, _D, _i64, _int, _I, _D, _D, _i64, _I, _I, _D, _D,
                                         _U8, _D, _int, _PP, _S]),
            "ml_thermal_add_loads": (_int, [_P, _i64, _I, _D, _i64, _I, _I, _D, _D, _D]),
            "ml_thermal_solve_step": (_int, [_P, C.c_double, _i64, C.c_double, _D, _D, _D, _D,
                                             _D, _S]),
            "ml_thermal_destroy": (None, [_P]),
            "ml_thermal_solve_steady": (_int, [_P, _D, _D, _D, _D, _S]),
            "ml_model_solve_large_strain": (_int, [_P, _U8, _D, _i64, _I, _I, _D, C.c_double,
                                                   _i64, _i64, _D, _D, _D, _D, _D, _D, _I,
                                                   C.POINTER(QuasiStaticInfo)]),
            "ml_model_solve_quasi_static": (_int, [_P, _D, _D, _D, C.c_double, _i64, _i64, _D,
                                                   _D, _D, _D, _D, _D, _D, _I,
                                                   C.POINTER(QuasiStaticInfo)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# Live handles, freed explicitly: close() (or a with-block) frees one, and at interpreter exit
# every handle still open is freed, newest first (a Batch before the Models it holds), before
# the interpreter tears down and while the HIP runtime and any tool attached to it are alive.
_live = weakref.WeakValueDictionary()
_serial = itertools.count()


class _Handle:
    """A library handle with deterministic release."""

    _destroy = ""  # name of the library function that frees the handle

    def _register(self, handle):
        self._h = handle
        self._key = next(_serial)
        _live[self._key] = self

    def close(self):
        """Frees the handle's device memory and stream now; later calls raise."""
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            getattr(_lib, self._destroy)(h)
        self._h = None
        if _live is not None:  # None once the interpreter has torn the module down
            _live.pop(getattr(self, "_key", None), None)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        self.close()


def close_all():
    """Frees every handle still open, newest first."""
    for key in sorted(list(_live.keys()), reverse=True):
        obj = _live.get(key)
        if obj is not None:
            obj.close()


atexit.register(close_all)


def _check(status, info=None):
    if status == ML_OK:
        return
    if status == ML_SINGULAR:
        raise Singular("singular matrix")
    if status == ML_NOT_CONVERGED:
        it = getattr(info, "iterations", getattr(info, "pcg_iterations", -1))
        res = getattr(info, "residual", float("nan"))
        raise IterativeNotConverged(it, res)
    if status in (ML_NEWTON_NOT_CONVERGED, ML_PLASTIC_COLLAPSE):
        inc = getattr(info, "failed_increment", -1)
        t = getattr(info, "failed_time", float("nan"))
        raise (NotConverged if status == ML_NEWTON_NOT_CONVERGED else PlasticCollapse)(inc, t)
    if status == ML_DIRECT_REQUIRED:
        raise DirectRequired(lib().ml_last_error().decode(errors="replace"))
    if status == ML_INVALID_MESH and info is not None:
        raise InvalidMesh(info.mesh_error, info.bad_element)
    msg = lib().ml_last_error().decode(errors="replace")
    if status == ML_DEVICE_TIMEOUT:
        raise DeviceError(f"device wait expired: {msg}")
    if status in (ML_INVALID_ARGUMENT, ML_OUT_OF_PATTERN):
        raise ValueError(msg or f"status {status}")
    raise DeviceError(f"status {status}: {msg}")


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ai64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def device_count() -> int:
    return lib().ml_device_count()


# ------------------------------------------------------------------ element level
def solid_d(young: float, nu: float) -> np.ndarray:
    """element.rs:193 — 6x6 isotropic D in Voigt order (xx, yy, zz, yz, xz, xy)."""
    d = np.zeros((6, 6))
    _check(lib().ml_solid_d(young, nu, _d(d)))
    return d


def stiffness_batch(coords: np.ndarray, young, nu) -> np.ndarray:
    """solid_stiffness (element.rs:223) on the device for a batch: coords (ne, nen, 3) ->
    Ke (ne, 3nen, 3nen) with Ke[e][j][i] as the reference's ``ke[j][i]``."""
    coords = _f64(coords)
    ne, nen, _ = coords.shape
    young = _f64(np.broadcast_to(young, (ne,)))
    nu = _f64(np.broadcast_to(nu, (ne,)))
    m = 3 * nen
    ke = np.zeros((ne, m, m))
    _check(lib().ml_stiffness_batch(nen, ne, _d(coords), _d(young), _d(nu), _d(ke)))
    return ke


def solid_stiffness(coords: np.ndarray, young: float, nu: float) -> np.ndarray:
    """solid_stiffness of one tetrahedron (Tet4: 4x3 coords, Tet10: 10x3)."""
    return stiffness_batch(np.asarray(coords)[None], young, nu)[0]


# ------------------------------------------------------------------ graph, RCM, CSR pattern
class Graph(_Handle):
    """iterative.rs:19-77 — DOF adjacency, built with its RCM order and lower pattern."""

    _destroy = "ml_pattern_free"

    def __init__(self, n, elements):
        lens = [len(e) for e in elements]
        off = np.zeros(len(elements) + 1, dtype=np.int64)
        off[1:] = np.cumsum(lens)
        items = np.array([v for e in elements for v in e] or [0], dtype=np.int64)
        h = C.c_void_p()
        _check(lib().ml_pattern_build(n, len(elements), _i(off), _i(items), C.byref(h)))
        self._register(h)
        self.n = n

    @classmethod
    def from_elements(cls, n, elements):
        return cls(n, elements)

    def neighbours(self, v):
        off, tgt = self.arrays()
        return tgt[off[v]:off[v + 1]]

    def arrays(self):
        nt = lib().ml_pattern_ntargets(self._h)
        off = np.zeros(self.n + 1, dtype=np.int64)
        tgt = np.zeros(max(nt, 1), dtype=np.int64)
        _check(lib().ml_pattern_graph(self._h, _i(off), _i(tgt)))
        return off, tgt[:nt]

    def reverse_cuthill_mckee(self):
        """solve.rs:200 — order[new] = old."""
        order = np.zeros(self.n, dtype=np.int64)
        _check(lib().ml_pattern_get(self._h, _i(order), None, None))
        return order

    def lower_pattern(self):
        """Csr::zeros in the RCM numbering (iterative.rs:91): offsets, cols."""
        nnz = lib().ml_pattern_nnz(self._h)
        off = np.zeros(self.n + 1, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int64)
        _check(lib().ml_pattern_get(self._h, None, _i(off), _i(cols)))
        return off, cols



def reverse_cuthill_mckee(graph: Graph) -> np.ndarray:
    return graph.reverse_cuthill_mckee()


class Csr:
    """iterative.rs:82-128 — host-side lower-triangle CSR (ascending columns, diagonal last)
    whose values are handed to the device by ``Pcg``."""

    def __init__(self, offsets, cols, values=None):
        self.offsets = _ai64(offsets)
        self.cols = _ai64(cols)
        self.values = np.zeros(len(self.cols)) if values is None else _f64(values).copy()

    @classmethod
    def zeros(cls, graph: Graph, order=None):
        """Csr::zeros (iterative.rs:91) in the graph's RCM numbering."""
        off, cols = graph.lower_pattern()
        return cls(off, cols)

    def n(self):
        return len(self.offsets) - 1

    def add(self, i, j, v):
        """Csr::add (iterative.rs:121): v into (i, j) and its mirror, inside the pattern."""
        if j > i:
            i, j = j, i
        row = slice(self.offsets[i], self.offsets[i + 1])
        k = np.searchsorted(self.cols[row], j)
        if k >= row.stop - row.start or self.cols[row][k] != j:
            raise ValueError("terme hors du profil")
        self.values[row.start + k] += v


def layout_verify(offsets, cols, l=None, r=None):
    """Builds and checks, on the host, every index array the device kernels would read for
    this lower CSR pattern (raises ValueError naming the failed check). Returns the layout
    statistics (block counts, largest external/internal block loads, DAG depth, longest row)
    and, when the factor values ``l`` and a vector ``r`` are given, z = (L L^T)^-1 r applied
    through the triangular-solve layouts in the device's order."""
    off, cols = _ai64(offsets), _ai64(cols)
    stats = np.zeros(8, dtype=np.int64)
    z = None
    if l is not None and r is not None:
        l, r = _f64(l), _f64(r)
        z = np.zeros(len(off) - 1)
    _check(lib().ml_layout_verify(len(off) - 1, _i(off), _i(cols),
                                  _d(l) if z is not None else None,
                                  _d(r) if z is not None else None,
                                  _d(z) if z is not None else None, _i(stats)))
    keys = ("fw_blocks", "bw_blocks", "fw_max_ext", "fw_max_int", "bw_max_ext", "bw_max_int",
            "depth", "max_row")
    out = dict(zip(keys, (int(v) for v in stats)))
    return (out, z) if z is not None else out


def _fused_code(fused):
    return int(fused) if not isinstance(fused, bool) else (1 if fused else 0)


def layout_sweeps(offsets, cols, l=None, r=None, sweeps=1, rows=256):
    """Builds and checks, on the host, the fused Jacobi-sweep layouts of this lower CSR
    pattern in blocks of ``rows`` (256 or 512) rows (ValueError when it does not fit them);
    returns their statistics (blocks, longest row, most external columns of a lower and of an
    upper block, the fused variant code a handle on this pattern starts with, 0 for one launch
    per sweep) and, given the factor values ``l`` and ``r``, the sweeps applied on the host
    through those layouts, with the factor rounded to FP32."""
    off, cols = _ai64(offsets), _ai64(cols)
    stats = np.zeros(5, dtype=np.int64)
    z = None
    if l is not None and r is not None:
        l, r = _f64(l), _f64(r)
        z = np.zeros(len(off) - 1)
    _check(lib().ml_layout_sweeps(len(off) - 1, _i(off), _i(cols), int(rows),
                                  _d(l) if z is not None else None,
                                  _d(r) if z is not None else None, int(sweeps),
                                  _d(z) if z is not None else None, _i(stats)))
    out = dict(zip(("blocks", "width", "max_ext_lower", "max_ext_upper", "default_variant"),
                   (int(v) for v in stats)))
    return (out, z) if z is not None else out


class Pcg(_Handle):
    """IC(0)-preconditioned CG on the device (iterative.rs:194-248). ``precision`` FP64 is the
    drop-in; MIXED runs an FP32 operator inside FP64 refinement."""

    _destroy = "ml_pcg_destroy"

    def __init__(self, a: Csr, precision: int = FP64):
        h = C.c_void_p()
        _check(lib().ml_pcg_create(a.n(), _i(a.offsets), _i(a.cols), _d(a.values), precision,
                                   C.byref(h)))
        self._register(h)
        self.n = a.n()
        self.nnz = len(a.cols)

    def solve(self, b, tol=0.0, maxit=0, info=False):
        """Pcg::solve: x with |r| <= 1e-12 |b| (or raises like the reference). b may hold
        several right-hand sides as rows of a 2-D array."""
        b = _f64(b)
        nrhs = 1 if b.ndim == 1 else b.shape[0]
        x = np.zeros_like(b)
        si = SolveInfo()
        _check(lib().ml_pcg_solve(self._h, _d(b), _d(x), nrhs, tol, maxit, C.byref(si)), si)
        return (x, si.as_dict()) if info else x

    def multiply(self, x):
        x = _f64(x)
        y = np.zeros(self.n)
        _check(lib().ml_pcg_multiply(self._h, _d(x), _d(y)))
        return y

    def precondition(self, r):
        r = _f64(r)
        z = np.zeros(self.n)
        _check(lib().ml_pcg_precondition(self._h, _d(r), _d(z)))
        return z

    def multiply_many(self, X):
        """A x for the rows of X (k x n) as batched solves multiply them."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.zeros_like(X)
        _check(lib().ml_pcg_multiply_many(self._h, X.shape[0], _d(X), _d(Y)))
        return Y

    def precondition_many(self, R):
        """The preconditioner applied to the rows of R (k x n) as batched solves apply it."""
        R = np.ascontiguousarray(R, dtype=np.float64)
        Z = np.zeros_like(R)
        _check(lib().ml_pcg_precondition_many(self._h, R.shape[0], _d(R), _d(Z)))
        return Z

    def factor(self):
        """(L values aligned with the CSR, shift index)."""
        l = np.zeros(self.nnz)
        sh = C.c_int32(-1)
        _check(lib().ml_pcg_factor(self._h, _d(l), C.byref(sh)))
        return l, sh.value

    def autotune(self, nrhs: int = 1):
        """Times candidate schedules and keeps the fastest; returns (waves per CU, right-hand
        sides per pass, reported occupancy per CU)."""
        out = (C.c_int32 * 3)()
        _check(lib().ml_pcg_autotune(self._h, int(nrhs), out))
        return tuple(out)

    def batch_lanes(self) -> int:
        """Lanes per row of the batched SpMV and sweeps (64 / rows per wave) the autotuner
        chose for several right-hand sides, or as set; 0: from the vectors of a pass."""
        return lib().ml_pcg_batch_lanes(self._h)

    def set_batch_lanes(self, lanes: int):
        """Lanes per row of the batched SpMV and sweeps: 8, 16, 32, 64, or 0 (from the pass)."""
        _check(lib().ml_pcg_set_batch_lanes(self._h, int(lanes)))

    def set_wait_limit(self, ticks: int = 0):
        """Bound of the device's dataflow waits in wall-clock ticks (0: the default, 2 s)."""
        _check(lib().ml_pcg_set_wait_limit(self._h, int(ticks)))

    def set_sweeps(self, sweeps: int):
        """Jacobi sweeps per triangular system with which a mixed handle applies its FP32
        factor (0: the exact triangular solves)."""
        _check(lib().ml_pcg_set_sweeps(self._h, int(sweeps)))

    def sweeps(self) -> int:
        return lib().ml_pcg_sweeps(self._h)

    def set_sweep_weight(self, w: float):
        """Weight of the sweeps (0 < w <= 1; 1: undamped)."""
        _check(lib().ml_pcg_set_sweep_weight(self._h, float(w)))

    def sweep_weight(self) -> float:
        return lib().ml_pcg_sweep_weight(self._h)

    def set_sweeps_fused(self, fused):
        """All sweeps of a triangular system in one launch (True, the default, where the
        pattern fits the fused kernel; or an int variant code 100 * rows per lane + register
        capacity: 116, 124, 132, 148, 180, 216, 224, 232) or one launch per sweep (False); the
        results are identical."""
        _check(lib().ml_pcg_set_sweeps_fused(self._h, _fused_code(fused)))

    def sweeps_fused(self) -> int:
        """Variant code of the fused kernel applying the sweeps (0: one launch each)."""
        return lib().ml_pcg_sweeps_fused(self._h)

    def set_reference_dots(self, on: bool = True):
        """PCG dot products in the reference's eight-lane order (solve.rs:100-110): with FP64,
        Pcg::solve is then the reference's bit for bit."""
        _check(lib().ml_pcg_set_reference_dots(self._h, int(bool(on))))



# ------------------------------------------------------------------ the static model
class Model(_Handle):
    """Model::new + Model::solve (model.rs:251-548) for a Tet4/Tet10 mesh on the device.
    ``problem`` is a ``maille_amd.meshes.Problem`` (nodes, conn, young, nu, fixed, imposed,
    nodal loads f, dead pressures). ``solver`` ITERATIVE (the default) always runs the device
    PCG, as ``with_solver(Iterative)``; AUTO raises DirectRequired below ITERATIVE_FROM free
    DOFs, where the reference's Auto is the direct solver."""

    _destroy = "ml_model_destroy"

    def __init__(self, problem, precision: int = FP64, solver: int = ITERATIVE):
        self.problem = problem
        nodes = _f64(problem.nodes)
        conn = _ai64(problem.conn)
        ne, nen = conn.shape
        young = _f64(np.broadcast_to(problem.young, (ne,)))
        nu = _f64(np.broadcast_to(problem.nu, (ne,)))
        fixed = np.ascontiguousarray(problem.fixed, dtype=np.uint8)
        imposed = _f64(problem.imposed)
        self.ndof = 3 * len(nodes)
        self.n_nodes = len(nodes)
        self.ne, self.nen = ne, nen
        h = C.c_void_p()
        si = SolveInfo()
        _check(lib().ml_model_create(len(nodes), _d(nodes), ne, nen, _i(conn), _d(young),
                                     _d(nu), fixed.ctypes.data_as(_U8), _d(imposed), solver,
                                     precision, C.byref(h), C.byref(si)), si)
        self._register(h)
        self.create_info = si.as_dict()

    def loads(self):
        """Model::assemble's F (model.rs:1089-1127): nodal loads, then dead pressures, then dead
        line loads on edges."""
        f = _f64(self.problem.f).copy()
        p = self.problem.pressures
        if p:
            arr = np.array(p, dtype=np.float64)
            e, fc, v = _ai64(arr[:, 0]), _ai64(arr[:, 1]), _f64(arr[:, 2])
            _check(lib().ml_model_add_pressures(self._h, len(e), _i(e), _i(fc), _d(v), _d(f)))
        el = getattr(self.problem, "edge_loads", None)
        if el:
            e = _ai64([l[0] for l in el])
            ed = _ai64([l[1] for l in el])
            q = _f64(np.array([l[2] for l in el], dtype=np.float64).reshape(-1, 3))
            _check(lib().ml_model_add_edge_loads(self._h, len(e), _i(e), _i(ed), _d(q), _d(f)))
        return f

    def solve(self, f=None, reactions=True):
        """(displacements over all DOFs, reactions or None, info dict)."""
        f = self.loads() if f is None else _f64(f)
        u = np.zeros(self.ndof)
        r = np.zeros(self.ndof) if reactions else None
        si = SolveInfo()
        _check(lib().ml_model_solve(self._h, _d(f), _d(u), _d(r) if reactions else None,
                                    C.byref(si)), si)
        return u, r, si.as_dict()

    def solve_device(self, f_ptr, u_ptr, r_ptr=None):
        """Solve with device buffers (integer device addresses of 3N doubles)."""
        si = SolveInfo()
        _check(lib().ml_model_solve_device(self._h, f_ptr, u_ptr, r_ptr, C.byref(si)), si)
        return si.as_dict()

    def factor(self):
        si = SolveInfo()
        _check(lib().ml_model_factor(self._h, C.byref(si)), si)
        return si.as_dict()

    def assemble(self):
        _check(lib().ml_model_assemble(self._h))

    def set_young(self, young):
        young = _f64(young)
        _check(lib().ml_model_set_young(self._h, _d(young)))

    def autotune(self, nrhs: int = 1):
        """Times candidate schedules and keeps the fastest; returns (waves per CU, right-hand
        sides per pass, reported occupancy per CU). A MIXED_JACOBI model also keeps the number
        of Jacobi sweeps whose whole solve was fastest (``sweeps()``)."""
        out = (C.c_int32 * 3)()
        _check(lib().ml_model_autotune(self._h, int(nrhs), out))
        return tuple(out)

    def batch_lanes(self) -> int:
        """Pcg.batch_lanes for the model's operator."""
        return lib().ml_model_batch_lanes(self._h)

    def set_batch_lanes(self, lanes: int):
        """Pcg.set_batch_lanes for the model's operator."""
        _check(lib().ml_model_set_batch_lanes(self._h, int(lanes)))

    def set_sweeps(self, sweeps: int):
        """Jacobi sweeps per triangular system for a mixed model's FP32 factor (0: exact)."""
        _check(lib().ml_model_set_sweeps(self._h, int(sweeps)))

    def sweeps(self) -> int:
        return lib().ml_model_sweeps(self._h)

    def set_sweep_weight(self, w: float):
        """Pcg.set_sweep_weight for the model's operator."""
        _check(lib().ml_model_set_sweep_weight(self._h, float(w)))

    def sweep_weight(self) -> float:
        return lib().ml_model_sweep_weight(self._h)

    def set_sweeps_fused(self, fused):
        """Pcg.set_sweeps_fused for the model's operator."""
        _check(lib().ml_model_set_sweeps_fused(self._h, _fused_code(fused)))

    def sweeps_fused(self) -> int:
        return lib().ml_model_sweeps_fused(self._h)

    def set_reference_dots(self, on: bool = True):
        """Pcg.set_reference_dots for the model's operator."""
        _check(lib().ml_model_set_reference_dots(self._h, int(bool(on))))

    def residency(self):
        """{kernel: (reported, used)} workgroups per CU of the dataflow kernels: the runtime's
        occupancy figure and the probed count the grids use (fused sweeps: zeros until their
        layouts are built)."""
        out = (C.c_int32 * 8)()
        _check(lib().ml_model_residency(self._h, out))
        return {k: (out[2 * i], out[2 * i + 1])
                for i, k in enumerate(("forward_solve", "backward_solve", "ic0", "fused_sweeps"))}

    def index_bytes(self):
        """(SpMV, sweeps) bytes per column index (2 or 4; sweeps 0 before their layout)."""
        out = (C.c_int32 * 2)()
        _check(lib().ml_model_index_bytes(self._h, out))
        return out[0], out[1]

    def spmv_index(self):
        """(form, bytes): the SpMV's column-index form (1: one 16-bit stream per group of rows
        with equal columns, 2: 16-bit offsets per entry, 4: int32) and the bytes of indices and
        per-row fields one SpMV reads in it."""
        out = np.zeros(2, dtype=np.int64)
        _check(lib().ml_model_spmv_index(self._h, _i(out)))
        return int(out[0]), int(out[1])

    def solve_many(self, F):
        """Load cases F (rows: nodal loads over all DOFs, without dead pressures) solved
        together; returns (U, list of info dicts)."""
        F = np.ascontiguousarray(np.atleast_2d(F), dtype=np.float64)
        k = F.shape[0]
        U = np.zeros_like(F)
        infos = (SolveInfo * k)()
        _check(lib().ml_model_solve_many(self._h, k, _d(F), _d(U), infos))
        return U, [i.as_dict() for i in infos]

    def compliance_gradient(self):
        """dC/dE_e of the compliance C = f.u at the last solution (adjoint: u itself)."""
        out = np.zeros(self.ne)
        _check(lib().ml_model_compliance_gradient(self._h, _d(out)))
        return out

    def stresses(self, u):
        u = _f64(u)
        ngp = 1 if self.nen == 4 else 4
        out = np.zeros((self.ne, ngp, 9))
        _check(lib().ml_model_stresses(self._h, _d(u), _d(out)))
        return out

    def stresses_device(self, u_ptr, out_ptr):
        """Gauss-point stresses with device buffers (integer device addresses)."""
        _check(lib().ml_model_stresses_device(self._h, u_ptr, out_ptr))

    def nodal_stresses(self, u):
        """results.rs:138-166 (nodal_stresses): per node x, y, z, then sxx, syy, szz, tyz, txz,
        txy averaged over the elements at the node, shape (n_nodes, 9)."""
        u = _f64(u)
        out = np.zeros((self.n_nodes, 9))
        _check(lib().ml_model_nodal_stresses(self._h, _d(u), _d(out)))
        return out

    def nfree(self):
        return lib().ml_model_nfree(self._h)

    def nnz(self):
        return lib().ml_model_nnz(self._h)

    def levels(self):
        return lib().ml_model_levels(self._h)

    def csr(self):
        """(order, offsets, cols, values) of K_ff in the reference layout."""
        n, nnz = self.nfree(), self.nnz()
        order = np.zeros(n, dtype=np.int64)
        off = np.zeros(n + 1, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int64)
        vals = np.zeros(nnz)
        _check(lib().ml_model_csr(self._h, _i(order), _i(off), _i(cols), _d(vals)))
        return order, off, cols, vals

    def profile(self, enable=-1):
        """Kernel durations (ms) and launch counts accumulated since the last reset; enable
        1/0 resets and switches the accounting on/off."""
        ms = np.zeros(10)
        cnt = np.zeros(10, dtype=np.int64)
        _check(lib().ml_model_profile(self._h, enable, _d(ms), _i(cnt)))
        names = ("spmv64", "spmv32", "precond64", "precond32", "ic0", "stiffness",
                 "assembly", "vector", "stresses", "nodal")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(names)}



class Batch(_Handle):
    """Several models solved in one set of launches (block-diagonal operator, one PCG per
    model). The models must stay alive and unchanged while the batch is used."""

    _destroy = "ml_batch_destroy"

    def __init__(self, models, precision: int = FP64):
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m._h.value for m in self.models])
        h = C.c_void_p()
        _check(lib().ml_batch_create(len(self.models), arr, precision, C.byref(h)))
        self._register(h)

    def solve(self, loads=None):
        """loads: one array of nodal loads per model (default: each model's loads());
        returns (list of displacements, list of info dicts)."""
        loads = [m.loads() for m in self.models] if loads is None else [_f64(f) for f in loads]
        us = [np.zeros(m.ndof) for m in self.models]
        fp = (_D * len(loads))(*[_d(f) for f in loads])
        up = (_D * len(us))(*[_d(u) for u in us])
        infos = (SolveInfo * len(us))()
        _check(lib().ml_batch_solve(self._h, fp, up, infos))
        return us, [i.as_dict() for i in infos]
