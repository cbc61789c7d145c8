"""The TOML model file (crates/maille-io/src/model of timothee-faget/maille v0.14.0) for the
static linear-elastic solve on tetrahedra.

``parse(text)`` reads the TOML and checks it against the schema (schema.rs: unknown keys and
wrong types refused, as ``deny_unknown_fields`` does); ``build(model, mesh)`` turns it and an
``MshFile`` into the ``Problem`` the device ``Model`` solves (build.rs:27-235, restricted to this
path); ``load(path)`` does both for a model file and the ``.msh`` it names (relative to the
model's folder, as the CLI opens it); ``solve_file(path)`` then solves it on the device.

On this path: ``analysis = "static"`` (or a model whose content makes it static, physics.rs),
small strains, no ``[step]``, tetrahedral meshes; materials with ``young`` and ``poisson``;
``sections``, ``bcs``, ``loads``, ``pressures`` and ``edge_loads``; ``[meshing]`` is accepted and
not used (it is the recipe for remeshing). Everything else the schema knows (thermal
analysis, ``heat_sources``, ``convections``, ``contacts``, ``temperatures``, ``step``,
``thermal_step``, ``kinematics = "large"``, plasticity, triangle meshes) is refused with a
``ModelError`` of kind ``"not_on_path"`` that names the key.

Errors are ``ModelError`` with ``kind`` naming the reference's variant (IoError::UnknownSet,
WrongSetKind, UnknownMaterial, UncoveredElements, OverlappingSections, InvalidModel, Toml, and
MailleError's DuplicateBc, MissingYoung, MissingPoisson, PoissonOutOfRange, InvalidProperty,
NonFinite) and the variant's fields as attributes.
"""
from __future__ import annotations

import math
import os

import numpy as np
import tomli

from . import msh as _msh
from .meshes import Problem


class ModelError(ValueError):
    def __init__(self, kind, message, **fields):
        super().__init__(message)
        self.kind = kind
        for k, v in fields.items():
            setattr(self, k, v)


def _invalid(message):
    return ModelError("invalid_model", message)


def _off_path(key, why):
    return ModelError("not_on_path", f"{key}: {why}; this library solves static linear "
                                     "elasticity on tetrahedra", key=key)


# ------------------------------------------------------------------ schema (schema.rs)
_STR, _NUM, _NUMS, _INT = "string", "number", "array of numbers", "integer"
_MATERIAL = {"young": _NUM, "poisson": _NUM, "conductivity": _NUM, "density": _NUM,
             "specific_heat": _NUM, "expansion": _NUM, "yield_stress": _NUM,
             "hardening": _NUM, "hyperelasticity": ("neo_hookean",)}
_SECTION = {"set": _STR, "material": _STR, "thickness": _NUM, "plane_state": ("stress", "strain")}
_VALUE = {"set": _STR, "dof": ("ux", "uy", "uz", "t"), "value": _NUM}
_PRESSURE = {"set": _STR, "value": _NUM}
_EDGE_LOAD = {"set": _STR, "value": _NUMS}
_REQUIRED = {"section": ("set", "material"), "value": ("set", "dof", "value"),
             "pressure": ("set", "value"), "edge_load": ("set", "value")}
_TOP = {"units": _STR, "mesh": _STR, "analysis": ("static", "thermal"),
        "kinematics": ("small", "large"), "materials": "materials", "sections": "sections",
        "bcs": "values", "loads": "values", "heat_sources": "tables",
        "convections": "tables", "pressures": "pressures", "edge_loads": "edge_loads",
        "contacts": "tables", "temperatures": "table", "step": "table", "meshing": "table",
        "thermal_step": "table"}


def _is_num(v):
    return isinstance(v, (int, float)) and not isinstance(v, bool)


def _check_value(where, v, kind):
    if kind == _STR:
        ok = isinstance(v, str)
    elif kind == _NUM:
        ok = _is_num(v)
    elif kind == _NUMS:
        ok = isinstance(v, list) and all(_is_num(x) for x in v)
    elif isinstance(kind, tuple):
        ok = isinstance(v, str) and v in kind
        if not ok:
            raise ModelError("toml", f"{where}: unknown variant {v!r}, expected one of "
                                     f"{', '.join(kind)}")
    else:
        ok = True
    if not ok:
        raise ModelError("toml", f"{where}: invalid type, expected {kind}")


def _check_table(where, table, schema, required=()):
    if not isinstance(table, dict):
        raise ModelError("toml", f"{where}: invalid type, expected a table")
    for key, v in table.items():
        if key not in schema:
            raise ModelError("toml", f"{where}: unknown field `{key}`, expected one of "
                                     f"{', '.join(schema)}")
        _check_value(f"{where}.{key}", v, schema[key])
    for key in required:
        if key not in table:
            raise ModelError("toml", f"{where}: missing field `{key}`")


def _check_list(where, items, schema, required):
    if not isinstance(items, list):
        raise ModelError("toml", f"{where}: invalid type, expected an array of tables")
    for i, t in enumerate(items):
        _check_table(f"{where}[{i}]", t, schema, required)


def parse(text: str) -> dict:
    """model::parse (mod.rs:20-23): the TOML, checked against ModelFile (schema.rs:49-100)."""
    try:
        doc = tomli.loads(text)
    except tomli.TOMLDecodeError as e:
        raise ModelError("toml", f"model file: {e}") from None
    for key, v in doc.items():
        if key not in _TOP:
            raise ModelError("toml", f"unknown field `{key}`, expected one of {', '.join(_TOP)}")
        kind = _TOP[key]
        if kind == "materials":
            if not isinstance(v, dict):
                raise ModelError("toml", "materials: invalid type, expected a table")
            for name, m in v.items():
                _check_table(f"materials.{name}", m, _MATERIAL)
        elif kind == "sections":
            _check_list(key, v, _SECTION, _REQUIRED["section"])
        elif kind == "values":
            _check_list(key, v, _VALUE, _REQUIRED["value"])
        elif kind == "pressures":
            _check_list(key, v, _PRESSURE, _REQUIRED["pressure"])
        elif kind == "edge_loads":
            _check_list(key, v, _EDGE_LOAD, _REQUIRED["edge_load"])
        elif kind == "tables":
            if not isinstance(v, list) or not all(isinstance(t, dict) for t in v):
                raise ModelError("toml", f"{key}: invalid type, expected an array of tables")
        elif kind == "table":
            if not isinstance(v, dict):
                raise ModelError("toml", f"{key}: invalid type, expected a table")
        else:
            _check_value(key, v, kind)
    for key in ("units", "mesh"):
        if key not in doc:
            raise ModelError("toml", f"missing field `{key}`")
    return doc


# ------------------------------------------------------------------ build (build.rs)
def _static_or_refuse(doc):
    """physics.rs:27-88 restricted to the static case: which computation the file asks for."""
    if doc.get("analysis") is not None:
        if "thermal_step" in doc:
            raise _invalid("`[thermal_step]` only serves the coupled computation: remove "
                           "`analysis`, or write `[step]` for a thermal computation alone")
        if doc["analysis"] == "thermal":
            raise _off_path("analysis", "the thermal analysis is another physics")
        return
    t_dof = lambda d: d["dof"] == "t"
    step = doc.get("step")
    thermal = (any(t_dof(d) for d in doc.get("bcs", [])) or any(t_dof(d) for d in doc.get("loads", []))
               or bool(doc.get("heat_sources")) or bool(doc.get("convections"))
               or "thermal_step" in doc or (step is not None and step.get("kind") == "transient"))
    mechanical = (any(not t_dof(d) for d in doc.get("bcs", []))
                  or any(not t_dof(d) for d in doc.get("loads", []))
                  or bool(doc.get("pressures")) or bool(doc.get("edge_loads"))
                  or bool(doc.get("contacts")) or "kinematics" in doc or "temperatures" in doc
                  or (step is not None and step.get("kind") == "quasi_static"))
    if not thermal and not mechanical:
        raise _invalid("the model does not say which computation to run: add a BC (`t` for "
                       "thermal, `ux`, `uy` or `uz` for static), or write `analysis`")
    if thermal:
        raise _off_path("analysis", "the model has a thermal part, another physics")


def _positive(name, value):
    """material.rs:3-9."""
    if not (math.isfinite(value) and value > 0.0):
        raise ModelError("invalid_property", f"{name} must be positive and finite, not {value}",
                         name=name, value=value)
    return float(value)


def _material(name, d):
    """build.rs:285-327: (young, poisson), either None when the file does not give it."""
    young, nu, k = d.get("young"), d.get("poisson"), d.get("conductivity")
    if young is not None and nu is not None:
        if not (-1.0 < nu < 0.5):  # material.rs:145-153
            raise ModelError("poisson_out_of_range", f"material '{name}': Poisson's ratio {nu} "
                                                     "is outside (-1, 0.5)", poisson=nu)
        out = (_positive("E (Young's modulus)", young), float(nu))
    elif young is not None:
        out = (_positive("E (Young's modulus)", young), None)
    elif nu is not None:
        raise _invalid(f"material '{name}': `poisson` without `young`")
    elif k is not None:
        _positive("k (thermal conductivity)", k)
        out = (None, None)
    else:
        raise _invalid(f"material '{name}': give `young` (static) or `conductivity` (thermal)")
    if young is not None and k is not None:
        _positive("k (thermal conductivity)", k)
    for key, label in (("density", "rho (density)"), ("specific_heat", "c (specific heat)"),
                       ("expansion", "alpha (expansion coefficient)")):
        if d.get(key) is not None:
            _positive(label, d[key])
    if d.get("yield_stress") is not None:
        raise _off_path(f"materials.{name}.yield_stress", "plasticity is solved by increments")
    if d.get("hardening") is not None:
        raise _invalid(f"material '{name}': `hardening` without `yield_stress`")
    if d.get("hyperelasticity") is not None:
        raise _invalid(f"material '{name}': `hyperelasticity` needs kinematics = \"large\" (the "
                       "law exists in large strains only)")
    return out


def _find(mesh, key, name):
    """build.rs:237-243."""
    if name not in mesh.sets:
        known = ", ".join(f"'{k}'" for k in mesh.sets)
        raise ModelError("unknown_set", f"{key}: the mesh has no set '{name}'; known sets: "
                                        f"{known}", key=key, name=name, known=list(mesh.sets))
    return mesh.sets[name]


def _expect(ok, name, expected):
    """build.rs:245-255."""
    if not ok:
        raise ModelError("wrong_set_kind", f"set '{name}' does not fit here: it must be a set "
                                           f"{expected}", name=name, expected=expected)


def _node_set(mesh, key, name):
    """build.rs:257-265."""
    s = _find(mesh, key, name)
    _expect(len(s.nodes) > 0, name, "of nodes (a boundary or a point)")
    return s


_DOF = {"ux": 0, "uy": 1, "uz": 2}


def _node_dof(key, dof):
    """build.rs:267-283 (static analysis, 3D)."""
    if dof == "t":
        raise _invalid(f"{key}: `t` is the temperature, reserved to the thermal analysis: in "
                       "statics, use `ux`, `uy` or `uz`")
    return _DOF[dof]


def _finite(name, value):
    """model.rs:1286-1292."""
    if not math.isfinite(value):
        raise ModelError("non_finite", f"{name} is not finite ({value})", name=name,
                         value=value)


def _elements(doc, mesh, materials):
    """build.rs:407-476: each cell's material through its one section."""
    ne = len(mesh.cells)
    section_of = np.full(ne, -1, dtype=np.int64)
    sections = doc.get("sections", [])
    for i, d in enumerate(sections):
        key = f"sections[{i}]"
        s = _find(mesh, key, d["set"])
        _expect(len(s.elements) > 0, d["set"], "of elements (the domain)")
        if d["material"] not in materials:
            known = ", ".join(f"'{k}'" for k in materials)
            raise ModelError("unknown_material", f"{key}: unknown material '{d['material']}'; "
                                                 f"defined materials: {known}", key=key,
                             name=d["material"], known=list(materials))
        if "thickness" in d or "plane_state" in d:
            raise _invalid(f"{key}: `thickness` and `plane_state` only serve triangles: remove "
                           "them for tetrahedra")
        taken = section_of[s.elements]
        if (taken >= 0).any():
            j = int(np.argmax(taken >= 0))
            raise ModelError("overlapping_sections",
                             f"element {int(s.elements[j])} is in sections[{int(taken[j])}] and "
                             f"sections[{i}]: one section per element",
                             element=int(s.elements[j]), first=int(taken[j]), second=i)
        section_of[s.elements] = i
    missing = np.flatnonzero(section_of < 0)
    if missing.size:
        raise ModelError("uncovered_elements",
                         f"{missing.size} element(s) without a section, element {missing[0]} "
                         "among them: add a [[sections]] covering the whole domain",
                         count=int(missing.size), first=int(missing[0]))
    young = np.full(ne, np.nan)
    nu = np.full(ne, np.nan)
    for i, d in enumerate(sections):
        e, p = materials[d["material"]]
        rows = section_of == i
        young[rows] = np.nan if e is None else e
        nu[rows] = np.nan if p is None else p
    return young, nu


def _check_materials(young, nu):
    """Model::with_analysis (model.rs:263-284): a static analysis needs E and nu on every
    element, checked in element order."""
    bad = np.flatnonzero(np.isnan(young) | np.isnan(nu))
    if bad.size:
        e = int(bad[0])
        if np.isnan(young[e]):
            raise ModelError("missing_young", f"element {e} has no Young's modulus", element=e)
        raise ModelError("missing_poisson", f"element {e} has no Poisson's ratio", element=e)


def build(doc: dict, mesh, name: str = "model") -> Problem:
    """build.rs:27-235 for the static linear-elastic path: the ``Problem`` of a parsed model
    and its ``MshFile``."""
    if not doc["units"].strip():
        raise _invalid("`units` is empty: write the coherent system of units used, for "
                       "example \"N, mm, MPa\"")
    _static_or_refuse(doc)
    for key, why in (("temperatures", "thermal expansion from a thermal computation is "
                                      "another physics"),
                     ("heat_sources", "heat sources belong to the thermal analysis"),
                     ("convections", "convection belongs to the thermal analysis"),
                     ("contacts", "contact is solved by increments"),
                     ("step", "stepped computations (transient, quasi-static) are other "
                              "solves")):
        if doc.get(key):
            raise _off_path(key, why)
    if doc.get("kinematics") == "large":
        raise _off_path("kinematics", "large strains are solved by increments")
    if mesh.dim != 3:
        raise _off_path("mesh", "a triangle mesh is a plane problem")

    materials = {}
    for mname in sorted(doc.get("materials", {}), key=lambda s: s.encode()):
        materials[mname] = _material(mname, doc["materials"][mname])
    young, nu = _elements(doc, mesh, materials)

    n = len(mesh.nodes)
    bcs, loads = [], []
    for i, d in enumerate(doc.get("bcs", [])):
        key = f"bcs[{i}]"
        dof = _node_dof(key, d["dof"])
        bcs.append((_node_set(mesh, key, d["set"]).nodes, dof, d))
    for i, d in enumerate(doc.get("loads", [])):
        key = f"loads[{i}]"
        dof = _node_dof(key, d["dof"])
        loads.append((_node_set(mesh, key, d["set"]).nodes, dof, d))

    # Model::with_analysis (model.rs:263-295): materials, then each BC (finite, not twice on a
    # DOF), then each load (finite)
    _check_materials(young, nu)
    fixed = np.zeros(3 * n, dtype=bool)
    imposed = np.zeros(3 * n)
    for nodes, dof, d in bcs:
        _finite("value imposed by a BC", d["value"])
        idx = 3 * nodes + dof
        if fixed[idx].any():
            j = int(np.argmax(fixed[idx]))
            raise ModelError("duplicate_bc", f"two BCs on node {int(nodes[j])}, dof "
                                             f"{d['dof']}", node=int(nodes[j]), dof=d["dof"])
        fixed[idx] = True
        imposed[idx] = float(d["value"])
    f = np.zeros(3 * n)
    for nodes, dof, d in loads:
        _finite("value of a Load", d["value"])
    for nodes, dof, d in loads:  # Model::assemble (model.rs:1089-1092): a value per node
        for v in nodes:
            f[3 * v + dof] += float(d["value"])

    pressures = []
    for i, d in enumerate(doc.get("pressures", [])):
        s = _find(mesh, f"pressures[{i}]", d["set"])
        _expect(len(s.faces) > 0, d["set"], "of faces (a boundary)")
        pressures += [(int(e), int(fc), float(d["value"])) for e, fc in s.faces]
    for _, _, v in pressures:  # Model::with_pressures (model.rs:431-455)
        _finite("p (pressure)", v)

    edge_loads = []
    for i, d in enumerate(doc.get("edge_loads", [])):
        key = f"edge_loads[{i}]"
        s = _find(mesh, key, d["set"])
        if len(d["value"]) != 3:
            raise _invalid(f"{key}: `value` must have 3 components ([fx, fy, fz]), not "
                           f"{len(d['value'])}")
        _expect(len(s.edges) > 0, d["set"], "of edges (a line of the .msh)")
        q = tuple(float(v) for v in d["value"])
        edge_loads += [(int(e), int(k), q) for e, k in s.edges]
    for _, _, q in edge_loads:  # Model::with_edge_loads (model.rs:461-487)
        for v in q:
            _finite("q (line load)", v)

    return Problem(name, np.ascontiguousarray(mesh.nodes, dtype=np.float64),
                   np.ascontiguousarray(mesh.cells, dtype=np.int64), young, nu, fixed, imposed,
                   f, pressures, {"units": doc["units"]}, edge_loads)


def load(path) -> Problem:
    """Reads a model file and the ``.msh`` its ``mesh`` names (relative to the model's folder),
    then builds the Problem."""
    with open(path, "rb") as fh:
        text = fh.read().decode("utf-8")
    doc = parse(text)
    mesh_path = os.path.join(os.path.dirname(os.path.abspath(path)), doc["mesh"])
    mesh = _msh.read_file(mesh_path)
    name = os.path.splitext(os.path.basename(path))[0]
    return build(doc, mesh, name)


def solve_file(path, precision=None):
    """``maille solve model.toml`` for the static path: (problem, displacements over all DOFs,
    reactions, solve info) from the device Model."""
    from . import FP64, Model
    problem = load(path)
    model = Model(problem, FP64 if precision is None else precision)
    u, r, info = model.solve()
    return problem, u, r, info
