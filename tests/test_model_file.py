"""The TOML model file on the static path: the reference's model tests (crates/maille-io/tests/
model.rs) restated for tetrahedra, solved here by the oracle (the device solve of the same
files is in test_gpu_parity.py), and the refusals with the variant the reference names."""
import os

import numpy as np
import pytest

import oracle
import oracle.msh as om
from maille_amd import meshes as M
from maille_amd import model_file as mf
from maille_amd import msh as lm

import msh_cases as cases

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# tests/model.rs:198-226
CUBE = """
units = "N, mm, MPa"
mesh = "cube.msh"
analysis = "static"
[materials.acier]
young = 200e3
poisson = 0.25
[[sections]]
set = "domaine"
material = "acier"
[[bcs]]
set = "x0"
dof = "ux"
value = 0.0
[[bcs]]
set = "x1"
dof = "ux"
value = 0.01
[[bcs]]
set = "y0"
dof = "uy"
value = 0.0
[[bcs]]
set = "z0"
dof = "uz"
value = 0.0
"""
ROLLERS = CUBE.replace('[[bcs]]\nset = "x1"\ndof = "ux"\nvalue = 0.01\n', "")


def golden_cube():
    return lm.read_file(os.path.join(GOLDEN, "cube.msh"))


def cube_with(extra_groups):
    """The reference's example cube rewritten with more groups: the point groups of the file
    plus ``extra_groups`` (name, dim, items) over its nodes."""
    g = golden_cube()
    groups = [("domaine", 3, None)]
    for name in ("x0", "x1", "y0", "z0"):
        groups.append((name, 0, g.sets[name].nodes.tolist()))
    text = cases.write_msh(g.nodes, g.cells, groups + extra_groups, seed=1, blocks=2)
    return lm.read(text)


def x1_faces(g):
    e, f = M.faces_on(g.nodes, g.cells, lambda p: p[:, 0] > 0.5)
    return [tuple(int(v) for v in g.cells[a][list(M.TET_FACES[b])]) for a, b in zip(e, f)]


def x1_rim(g):
    """tests/model.rs:727-752: the four edges of the face x = 1 along y or z."""
    rim = set()
    for c in g.cells:
        for a, b in M.TET10_EDGES:
            pa, pb = g.nodes[c[a]], g.nodes[c[b]]
            if pa[0] > 0.5 and pb[0] > 0.5 and np.count_nonzero(pa != pb) == 1:
                rim.add(tuple(sorted((int(c[a]), int(c[b])))))
    assert len(rim) == 4
    return sorted(rim)


def solve(problem):
    r = oracle.model_solve(problem, iterative=False)
    assert r["status"] == oracle.OK
    return r


# ------------------------------------------------------------------ solved models
def test_cube_in_traction_matches_the_analytic_field():
    """tests/model.rs:194-228: eps = 0.01, ux = eps x, uy = -nu eps y, uz = -nu eps z."""
    g = golden_cube()
    p = mf.build(mf.parse(CUBE), g)
    assert p.meta["units"] == "N, mm, MPa"
    u = solve(p)["displacements"].reshape(-1, 3)
    x = g.nodes
    assert np.abs(u[:, 0] - 0.01 * x[:, 0]).max() < 1e-12
    assert np.abs(u[:, 1] + 0.25 * 0.01 * x[:, 1]).max() < 1e-12
    assert np.abs(u[:, 2] + 0.25 * 0.01 * x[:, 2]).max() < 1e-12


def test_pressure_on_a_face_set_pulls():
    """tests/model.rs:642-661 in 3D: a pressure of -100 on the face x = 1 over rollers is a
    uniaxial stress of 100, ux = 100 x / E, uy = -nu 100 y / E, uz = -nu 100 z / E."""
    g = cube_with([])
    g = cube_with([("bout", 2, x1_faces(g))])
    assert len(g.sets["bout"].faces) == 2
    p = mf.build(mf.parse(ROLLERS + '[[pressures]]\nset = "bout"\nvalue = -100.0\n'), g)
    assert len(p.pressures) == 2
    u = solve(p)["displacements"].reshape(-1, 3)
    e, nu, s = 200e3, 0.25, 100.0
    x = g.nodes
    assert np.abs(u[:, 0] - s / e * x[:, 0]).max() < 1e-12
    assert np.abs(u[:, 1] + nu * s / e * x[:, 1]).max() < 1e-12
    assert np.abs(u[:, 2] + nu * s / e * x[:, 2]).max() < 1e-12


def test_edge_load_on_an_edge_set_gives_the_total_force_in_3d():
    """tests/model.rs:754-801: q = 25 N/mm on the four edges of x = 1 (perimeter 4): the x
    reactions sum to -100 N; a node set is no edge set; the value needs 3 components."""
    g = cube_with([])
    g = cube_with([("contour", 1, x1_rim(g))])
    assert len(g.sets["contour"].edges) == 4
    toml = ROLLERS + '[[edge_loads]]\nset = "contour"\nvalue = [25.0, 0.0, 0.0]\n'
    r = solve(mf.build(mf.parse(toml), g))
    left = g.nodes[:, 0] < 0.5
    assert abs(r["reactions"][0::3][left].sum() + 100.0) < 1e-9
    with pytest.raises(mf.ModelError) as e:
        mf.build(mf.parse(toml.replace('set = "contour"', 'set = "x1"')), g)
    assert e.value.kind == "wrong_set_kind"
    with pytest.raises(mf.ModelError) as e:
        mf.build(mf.parse(toml.replace("[25.0, 0.0, 0.0]", "[25.0, 0.0]")), g)
    assert "3 components" in str(e.value)


def test_nodal_loads_are_per_node():
    """schema.rs:182-189: a [[loads]] value is the force on each node of the set, not the
    total: 25 N on each of the four nodes of x = 1 gives reactions summing to -100 N."""
    g = golden_cube()
    p = mf.build(mf.parse(ROLLERS + '[[loads]]\nset = "x1"\ndof = "ux"\nvalue = 25.0\n'), g)
    assert p.f[0::3].sum() == 100.0 and np.count_nonzero(p.f) == 4
    r = solve(p)
    assert abs(r["reactions"][0::3][g.nodes[:, 0] < 0.5].sum() + 100.0) < 1e-9


def test_load_reads_the_mesh_next_to_the_model(tmp_path):
    """The CLI opens `mesh` relative to the model's folder (maille-cli/src/mesh.rs:120)."""
    (tmp_path / "cube.msh").write_bytes(open(os.path.join(GOLDEN, "cube.msh"), "rb").read())
    (tmp_path / "cube.toml").write_text(CUBE)
    p = mf.load(str(tmp_path / "cube.toml"))
    q = mf.build(mf.parse(CUBE), golden_cube())
    assert p.name == "cube"
    for a in ("nodes", "conn", "young", "nu", "fixed", "imposed", "f"):
        assert np.array_equal(getattr(p, a), getattr(q, a)), a


def test_library_and_oracle_readers_build_the_same_problem():
    """The Problem built from the library's reading equals the one built from the oracle's."""
    g = cube_with([])
    text = cases.write_msh(g.nodes, g.cells, [("domaine", 3, None), ("bout", 2, x1_faces(g)),
                                              ("x0", 0, g.sets["x0"].nodes.tolist()),
                                              ("y0", 0, g.sets["y0"].nodes.tolist()),
                                              ("z0", 0, g.sets["z0"].nodes.tolist())], seed=5)
    o = om.read(text)
    as_lib = lm.MshFile(np.array(o.nodes), o.kind, np.array(o.cells, dtype=np.int64),
                        {k: lm.MshSet(np.array(s.nodes, dtype=np.int64),
                                      np.array(s.faces, dtype=np.int64).reshape(-1, 2),
                                      np.array(s.edges, dtype=np.int64).reshape(-1, 2),
                                      np.array(s.elements, dtype=np.int64))
                         for k, s in o.sets.items()})
    doc = mf.parse(ROLLERS + '[[pressures]]\nset = "bout"\nvalue = 3.0\n')
    a, b = mf.build(doc, lm.read(text)), mf.build(doc, as_lib)
    for attr in ("nodes", "conn", "young", "nu", "fixed", "imposed", "f"):
        assert np.array_equal(getattr(a, attr), getattr(b, attr)), attr
    assert a.pressures == b.pressures


# ------------------------------------------------------------------ refusals
def error(toml, mesh=None):
    with pytest.raises(mf.ModelError) as e:
        mf.build(mf.parse(toml), golden_cube() if mesh is None else mesh)
    return e.value


def test_unknown_key_and_wrong_type_are_refused():
    """tests/model.rs:465-472 (deny_unknown_fields)."""
    assert error(CUBE + "\nfoo = 1\n").kind == "toml"
    assert error(CUBE.replace("young = 200e3", "young = 200e3\nyoungs = 1.0")).kind == "toml"
    assert error(CUBE.replace('value = 0.01', 'value = "0.01"')).kind == "toml"
    assert error(CUBE.replace('dof = "uy"', 'dof = "uw"')).kind == "toml"
    assert error(CUBE.replace('units = "N, mm, MPa"\n', "")).kind == "toml"
    assert error("units = [\n").kind == "toml"


def test_units_are_required():
    """tests/model.rs:474-480."""
    e = error(CUBE.replace('"N, mm, MPa"', '"  "'))
    assert e.kind == "invalid_model" and "units" in str(e)


def test_unknown_set_lists_the_known_ones():
    """tests/model.rs:482-493."""
    e = error(CUBE.replace('set = "y0"', 'set = "dessus"'))
    assert e.kind == "unknown_set" and e.key == "bcs[2]" and e.name == "dessus"
    assert e.known == ["domaine", "x0", "x1", "y0", "z0"]
    e = error(CUBE + '[[pressures]]\nset = "dessus"\nvalue = 1.0\n')
    assert e.kind == "unknown_set" and e.key == "pressures[0]"
    e = error(CUBE + '[[edge_loads]]\nset = "dessus"\nvalue = [1.0, 0.0, 0.0]\n')
    assert e.kind == "unknown_set" and e.key == "edge_loads[0]"


def test_wrong_set_kind_is_refused():
    """tests/model.rs:495-515, 664-681, 703-725."""
    assert error(CUBE.replace('set = "domaine"', 'set = "x0"')).kind == "wrong_set_kind"
    assert error(CUBE.replace('set = "x0"', 'set = "domaine"')).kind == "wrong_set_kind"
    assert error(CUBE + '[[pressures]]\nset = "domaine"\nvalue = 1.0\n').kind == "wrong_set_kind"
    assert error(CUBE + '[[pressures]]\nset = "x1"\nvalue = 1.0\n').kind == "wrong_set_kind"
    e = error(CUBE + '[[edge_loads]]\nset = "domaine"\nvalue = [1.0, 0.0, 0.0]\n')
    assert e.kind == "wrong_set_kind"


def test_unknown_material_lists_the_defined_ones():
    """tests/model.rs:517-525."""
    e = error(CUBE.replace('material = "acier"', 'material = "alu"'))
    assert e.kind == "unknown_material" and e.name == "alu" and "'acier'" in str(e)


def test_sections_must_cover_every_element_once():
    """tests/model.rs:527-554."""
    g = golden_cube()
    g.sets["domaine"].elements = g.sets["domaine"].elements[:-1]
    e = error(CUBE, g)
    assert (e.kind, e.count, e.first) == ("uncovered_elements", 1, 5)
    e = error(CUBE + '[[sections]]\nset = "domaine"\nmaterial = "acier"\n')
    assert (e.kind, e.element, e.first, e.second) == ("overlapping_sections", 0, 0, 1)


def test_dof_must_fit_the_static_analysis():
    """tests/model.rs:556-564: `t` is the temperature."""
    e = error(CUBE.replace('dof = "uy"', 'dof = "t"'))
    assert e.kind == "invalid_model" and "thermal" in str(e)


def test_section_keys_must_fit_tetrahedra():
    """tests/model.rs:566-584."""
    e = error(CUBE.replace('material = "acier"\n', 'material = "acier"\nthickness = 1.0\n'))
    assert e.kind == "invalid_model" and "tetrahedra" in str(e)
    e = error(CUBE.replace('material = "acier"\n', 'material = "acier"\nplane_state = "strain"\n'))
    assert e.kind == "invalid_model"


def test_incomplete_or_invalid_materials_are_refused():
    """tests/model.rs:586-597, material.rs:3-9 and 145-153, model.rs:263-284."""
    assert "young" in str(error(CUBE.replace("young = 200e3\npoisson = 0.25", "density = 1.0")))
    assert "poisson" in str(error(CUBE.replace("young = 200e3\n", "")))
    assert "yield_stress" in str(error(CUBE.replace("poisson = 0.25", "poisson = 0.25\n"
                                                                      "hardening = 5.0")))
    assert error(CUBE.replace("poisson = 0.25", "poisson = 0.5")).kind == "poisson_out_of_range"
    assert error(CUBE.replace("young = 200e3", "young = -1.0")).kind == "invalid_property"
    assert error(CUBE.replace("poisson = 0.25", "poisson = 0.25\ndensity = 0.0")).kind == \
        "invalid_property"
    e = error(CUBE.replace("poisson = 0.25", ""))
    assert (e.kind, e.element) == ("missing_poisson", 0)
    e = error(CUBE.replace("young = 200e3\npoisson = 0.25", "conductivity = 1.0"))
    assert (e.kind, e.element) == ("missing_young", 0)


def test_bc_and_load_values_are_checked():
    """model.rs:285-299: finite values, one BC per DOF."""
    assert error(CUBE.replace("value = 0.01", "value = nan")).kind == "non_finite"
    e = error(CUBE + '[[bcs]]\nset = "x0"\ndof = "uy"\nvalue = 0.0\n')
    assert (e.kind, e.node, e.dof) == ("duplicate_bc", 0, "uy")
    assert error(CUBE + '[[loads]]\nset = "x1"\ndof = "uz"\nvalue = inf\n').kind == "non_finite"


@pytest.mark.parametrize("extra,key", [
    ('analysis = "thermal"', "analysis"),
    ('kinematics = "large"', "kinematics"),
    ('[[contacts]]\nset = "x1"\nkind = "plane"', "contacts"),
    ('[step]\nkind = "quasi_static"\nduration = 1.0\nincrements = 2', "step"),
    ('[temperatures]\nreference = 20.0', "temperatures"),
    ('[[heat_sources]]\nset = "domaine"\nvalue = 1.0', "heat_sources"),
])
def test_other_physics_is_refused_by_name(extra, key):
    """What the schema knows beyond static linear elasticity is refused, naming the key."""
    toml = CUBE.replace('analysis = "static"\n', "") if key == "analysis" else CUBE
    toml = (extra + "\n" + toml) if "=" in extra.split("\n")[0] and "[" not in extra else \
        toml + "\n" + extra + "\n"
    e = error(toml)
    assert e.kind == "not_on_path" and e.key == key, str(e)


def test_yield_stress_and_triangles_are_refused_by_name():
    e = error(CUBE.replace("poisson = 0.25", "poisson = 0.25\nyield_stress = 250.0"))
    assert e.kind == "not_on_path"
    e = error(CUBE, lm.read(cases.TRIANGLE))
    assert (e.kind, e.key) == ("not_on_path", "mesh")


def test_a_model_without_a_computation_is_refused():
    """physics.rs:60-63."""
    toml = CUBE.replace('analysis = "static"\n', "")
    toml = toml[:toml.index("[[bcs]]")]
    assert error(toml).kind == "invalid_model"
