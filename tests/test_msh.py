"""The Gmsh 4.1 reader: the reference's reader tests (crates/maille-io/tests/msh.rs and the
unit test of numbering.rs) restated and run on both the oracle's restatement (oracle/msh.py)
and the library's native reader (ml_msh_read, host code: no GPU needed), the reference's
example file examples/cube.msh (tests/golden/cube.msh), generated meshes whose reading is known
by construction, and refusal parity (same kind, same line) between the two readers on
corrupted files."""
import os

import numpy as np
import pytest

import oracle.msh as om
from maille_amd import meshes as M
from maille_amd import msh as lm

import msh_cases as cases

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _norm_oracle(f):
    sets = {k: (list(s.nodes), [tuple(x) for x in s.faces], [tuple(x) for x in s.edges],
                list(s.elements)) for k, s in f.sets.items()}
    return (np.array(f.nodes, dtype=float).reshape(-1, 3), f.kind,
            np.array(f.cells, dtype=np.int64).reshape(len(f.cells), -1), sets)


def _norm_library(f):
    sets = {k: (s.nodes.tolist(), [tuple(x) for x in s.faces.tolist()],
                [tuple(x) for x in s.edges.tolist()], s.elements.tolist())
            for k, s in f.sets.items()}
    return f.nodes, f.kind, f.cells, sets


READERS = {"oracle": lambda t: _norm_oracle(om.read(t)),
           "library": lambda t: _norm_library(lm.read(t))}
ERRORS = {"oracle": om.MshError, "library": lm.MshError}


def outcome(reader, text):
    try:
        return ("ok",) + READERS[reader](text)
    except ERRORS[reader] as e:
        return ("error", e.kind, e.line)


def same(a, b):
    if a[0] != b[0]:
        return False
    if a[0] == "error":
        return a[1:] == b[1:]
    return (a[1].tobytes() == b[1].tobytes() and a[2] == b[2] and np.array_equal(a[3], b[3])
            and a[4] == b[4])


def tet_volumes(nodes, cells):
    x = nodes[cells[:, :4]]
    return np.einsum("ij,ij->i", x[:, 1] - x[:, 0],
                     np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 0])) / 6


# ------------------------------------------------------------------ tests/msh.rs, restated
@pytest.mark.parametrize("reader", READERS)
def test_tet10_gmsh_numbering_and_orientation(reader):
    """tests/msh.rs:98-131: each mid-node at the middle of its edge in Maille's numbering, a
    quadratic field interpolated exactly, and both tetrahedra positive (the second is written
    reversed)."""
    nodes, kind, cells, _ = READERS[reader](cases.tet10_file())
    assert kind == "tet10" and cells.shape == (2, 10)
    quadratic = lambda p: 1 + p[0] + 2 * p[1] - p[2] + p[0] * p[1] + 3 * p[2] ** 2
    xi, eta, zeta = 0.2, 0.3, 0.1
    lam = [1 - xi - eta - zeta, xi, eta, zeta]
    shape = [l * (2 * l - 1) for l in lam] + [4 * lam[a] * lam[b] for a, b in M.TET10_EDGES]
    for c in cells:
        pt = nodes[c]
        for k, (a, b) in enumerate(M.TET10_EDGES):
            assert np.abs(pt[4 + k] - (pt[a] + pt[b]) / 2).max() < 1e-14
        x = sum(s * p for s, p in zip(shape, pt))
        u = sum(s * quadratic(p) for s, p in zip(shape, pt))
        assert abs(u - quadratic(x)) < 1e-12
    assert (tet_volumes(nodes, cells) > 0).all()


@pytest.mark.parametrize("reader", READERS)
def test_tet10_sets(reader):
    """tests/msh.rs:133-148."""
    nodes, kind, cells, sets = READERS[reader](cases.tet10_file())
    assert sets["corps"][3] == [0, 1]
    assert sets["coin"][0] == [0]
    assert sets["base"][1] == [(0, 0)]
    assert sets["base"][0] == sorted(int(cells[0][p]) for p in (0, 1, 2, 4, 5, 6))


@pytest.mark.parametrize("reader", READERS)
def test_gmsh_line_group_becomes_an_edge_set(reader):
    """tests/msh.rs:171-184: line 1-4 (mid-node 8) is edge 3 (0-3) of the first tetrahedron."""
    nodes, kind, cells, sets = READERS[reader](cases.tet10_file_with_line("1 4 8"))
    arete = sets["arete"]
    assert arete[2] == [(0, 3)]
    assert arete[0] == sorted(int(cells[0][p]) for p in (0, 3, 7))
    assert arete[1] == [] and arete[3] == []
    assert sets["coin"][0] == [0] and sets["base"][1] == [(0, 0)]


@pytest.mark.parametrize("reader", READERS)
def test_refuses_line_that_is_not_an_edge(reader):
    """tests/msh.rs:224-231: nodes 1 and 11 are vertices of two different tetrahedra."""
    r = outcome(reader, cases.tet10_file_with_line("1 11 5"))
    assert r[:2] == ("error", "boundary_edge_not_found")
    if reader == "oracle":
        with pytest.raises(om.MshError) as e:
            om.read(cases.tet10_file_with_line("1 11 5"))
        assert e.value.set == "arete"


@pytest.mark.parametrize("reader", READERS)
def test_clockwise_triangle_is_flipped(reader):
    """tests/msh.rs:284-292."""
    nodes, kind, cells, sets = READERS[reader](cases.TRIANGLE.replace("1 1 2 3", "1 1 3 2"))
    p = nodes[cells[0]]
    area2 = (p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1])
    assert kind == "tri3" and area2 > 0
    assert sets["plaque"][3] == [0]


def test_reversed_triangle_is_flipped_by_orient():
    """numbering.rs:164-177."""
    pts = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)]
    ccw, cw = [0, 1, 2], [0, 2, 1]
    om.orient("tri3", ccw, pts)
    om.orient("tri3", cw, pts)
    assert ccw == [0, 1, 2] and cw == [2, 0, 1]


@pytest.mark.parametrize("reader", READERS)
def test_unnamed_group_is_named_by_its_number(reader):
    """tests/msh.rs:294-301."""
    text = (cases.TRIANGLE.replace('2 1 "plaque"', "").replace("$PhysicalNames\n1\n", "")
            .replace("$EndPhysicalNames\n", ""))
    assert READERS[reader](text)[3]["1"][3] == [0]


REFUSALS = {
    # tests/msh.rs:307-397: (text, kind, line)
    "old_version": (cases.TRIANGLE.replace("4.1 0 8", "2.2 0 8"), "unsupported_version", 2),
    "binary": (cases.TRIANGLE.replace("4.1 0 8", "4.1 1 8"), "binary", 2),
    "hexahedra": (cases.TRIANGLE.replace("2 1 2 1\n1 1 2 3", "3 1 5 1\n1 1 2 3 1 2 3 1 2"),
                  "unsupported_element", 24),
    "mixed_orders": (cases.TRIANGLE.replace("1 1 1 1\n2 1 2 1\n1 1 2 3",
                                            "2 2 1 2\n2 1 2 1\n1 1 2 3\n2 1 9 1\n2 1 2 3 1 2 3"),
                     "mixed_cell_kinds", None),
    "no_cells": (cases.TRIANGLE.replace("1 1 1 1\n2 1 2 1\n1 1 2 3", "0 0 0 0"), "no_cells",
                 None),
    "missing_section": (cases.TRIANGLE.replace("$Elements", "$Autre")
                        .replace("$EndElements", "$EndAutre"), "missing_section", None),
    "nan": (cases.TRIANGLE.replace("\n1 0 0\n0 1 0", "\nnan 0 0\n0 1 0"), "syntax", 19),
    "inf": (cases.TRIANGLE.replace("\n1 0 0\n0 1 0", "\ninf 0 0\n0 1 0"), "syntax", 19),
    "-inf": (cases.TRIANGLE.replace("\n1 0 0\n0 1 0", "\n-inf 0 0\n0 1 0"), "syntax", 19),
    "unclosed": (cases.TRIANGLE.replace("$EndElements\n", ""), "syntax", 25),
    "unknown_node": (cases.TRIANGLE.replace("1 1 2 3", "1 1 2 9"), "unknown_node", 25),
    "non_planar": (cases.TRIANGLE.replace("0 1 0\n$EndNodes", "0 1 0.5\n$EndNodes"),
                   "non_planar", None),
    "bad_number": (cases.TRIANGLE.replace("1 0 0\n", "1 zéro 0\n"), "syntax", 19),
    "duplicate_group_name": (
        cases.TRIANGLE.replace('1\n2 1 "plaque"', '2\n2 1 "x"\n1 2 "x"')
        .replace("0 0 1 0\n", "0 1 1 0\n1 0 0 0 1 0 0 1 2 0\n")
        .replace("1 1 1 1\n2 1 2 1", "2 2 1 2\n1 1 1 1\n2 1 2\n2 1 2 1"),
        "duplicate_set_name", None),
    "boundary_not_a_face": (
        cases.TRIANGLE.replace('2 1 "plaque"', '1 1 "bord"')
        .replace("0 0 1 0\n1 0 0 0 1 1 0 1 1 0", "0 1 0 0\n1 0 0 0 1 0 0 1 1 0")
        .replace("1 3 1 3\n2 1 0 3\n1\n2\n3\n", "1 4 1 4\n2 1 0 4\n1\n2\n3\n4\n")
        .replace("0 1 0\n$EndNodes", "0 1 0\n2 2 0\n$EndNodes")
        .replace("1 1 1 1\n2 1 2 1\n1 1 2 3", "2 2 1 2\n2 1 2 1\n1 1 2 3\n1 1 1 1\n2 1 4"),
        "boundary_face_not_found", None),
}


@pytest.mark.parametrize("reader", READERS)
@pytest.mark.parametrize("case", REFUSALS)
def test_refusals(reader, case):
    """tests/msh.rs:307-397: the variant, and the line where the reference's test names it (or
    where the reference's reader places it: the header of the unsupported block, the last line
    of an unclosed section)."""
    text, kind, line = REFUSALS[case]
    r = outcome(reader, text)
    assert r[0] == "error" and r[1] == kind, r
    if line is not None:
        assert r[2] == line, r


def test_refusal_messages_name_what_to_do():
    """tests/msh.rs:314, 324, 333, 367-371: the messages say how to fix the file."""
    def msg(text):
        with pytest.raises(lm.MshError) as e:
            lm.read(text)
        return str(e.value)
    assert "Mesh.MshFileVersion = 4.1" in msg(REFUSALS["old_version"][0])
    assert "Mesh.Binary = 0" in msg(REFUSALS["binary"][0])
    assert "hexahedron" in msg(REFUSALS["hexahedra"][0])
    m = msg(REFUSALS["nan"][0])
    assert "line 19" in m and "node 2" in m and "not finite" in m


# ------------------------------------------------------------------ the reference's example
def test_reference_example_cube():
    """examples/cube.msh: six Tet4 of the unit cube (elements 10, 12 and 14 of the file are
    written with negative volume), point groups on the faces x = 0, x = 1, y = 0, z = 0 and
    the volume group "domaine"."""
    path = os.path.join(GOLDEN, "cube.msh")
    text = open(path).read()
    a = _norm_library(lm.read_file(path))
    b = _norm_oracle(om.read(text))
    assert same(("ok",) + a, ("ok",) + b)
    nodes, kind, cells, sets = a
    assert kind == "tet4" and cells.shape == (6, 4) and len(nodes) == 8
    assert np.allclose(tet_volumes(nodes, cells), 1 / 6, rtol=0, atol=1e-15)
    # vertex sets as written in the file (Gmsh tags 1..8 are indices 0..7)
    written = [(1, 2, 4, 8), (1, 2, 6, 8), (1, 3, 4, 8), (1, 3, 7, 8), (1, 5, 6, 8), (1, 5, 7, 8)]
    assert [sorted(c) for c in cells.tolist()] == [[t - 1 for t in w] for w in written]
    for name, axis, value in (("x0", 0, 0), ("x1", 0, 1), ("y0", 1, 0), ("z0", 2, 0)):
        assert sets[name][0] == np.flatnonzero(nodes[:, axis] == value).tolist()
    assert sets["domaine"][3] == list(range(6))
    assert list(sets) == ["domaine", "x0", "x1", "y0", "z0"]


# ------------------------------------------------------------------ generated meshes
def _box_case(quadratic, seed, sparse_tags):
    nodes, tets = M.box_grid((4, 3, 2), (2.0, 1.5, 1.0))
    cells = M.to_quadratic_tet(nodes, tets)[1] if quadratic else tets
    if quadratic:
        nodes = M.to_quadratic_tet(nodes, tets)[0]
    assert (tet_volumes(nodes, cells) > 0).all()
    x = nodes
    on = lambda pred: [tuple(map(int, cells[e][list(M.TET_FACES[f])]))
                       for e, f in zip(*M.faces_on(x, cells, pred))]
    left = on(lambda p: np.abs(p[..., 0]) < 1e-12)
    rim = []
    for c in cells:
        for a, b in M.TET10_EDGES:
            pa, pb = x[c[a]], x[c[b]]
            if abs(pa[0] - 2) < 1e-12 and abs(pb[0] - 2) < 1e-12 and \
                    (np.count_nonzero(np.abs(pa - pb) > 1e-12) == 1):
                rim.append((int(c[a]), int(c[b])))
    rim = sorted(set(tuple(sorted(r)) for r in rim))
    corner = [int(np.argmin(np.abs(x).sum(axis=1)))]
    groups = [("solide", 3, None), ("gauche", 2, left), ("bord", 1, rim), ("coin", 0, corner)]
    text = cases.write_msh(x, cells, groups, seed=seed, sparse_tags=sparse_tags)
    return x, cells, groups, text


@pytest.mark.parametrize("quadratic", [False, True])
@pytest.mark.parametrize("sparse_tags", [False, True])
def test_generated_mesh_reads_back_exactly(quadratic, sparse_tags):
    """A box written with shuffled (or sparse) tags, two node blocks, three volume blocks and
    half the cells reversed: both readers give back the cells and nodes bit for bit, and the
    sets the construction implies (faces and edges at their first element, found here by brute
    force)."""
    x, cells, groups, text = _box_case(quadratic, seed=7, sparse_tags=sparse_tags)
    a = ("ok",) + _norm_library(lm.read(text))
    b = ("ok",) + _norm_oracle(om.read(text))
    assert same(a, b)
    nodes, kind, got, sets = a[1:]
    assert kind == ("tet10" if quadratic else "tet4")
    assert nodes.tobytes() == x.tobytes() and np.array_equal(got, cells)
    nen = cells.shape[1]
    assert sets["solide"][3] == list(range(len(cells)))
    fpos = [om.face_positions("tet10" if quadratic else "tet4", f) for f in range(4)]
    owner = cases.first_owner(cells, fpos, 3, groups[1][2])
    assert sets["gauche"][1] == sorted(owner.values())
    assert sets["gauche"][0] == sorted({int(cells[e][p]) for e, f in owner.values()
                                        for p in fpos[f]})
    epos = [om.edge_positions("tet10" if quadratic else "tet4", k) for k in range(6)]
    owner = cases.first_owner(cells, epos, 2, groups[2][2])
    assert len(owner) == len(groups[2][2]) and sets["bord"][2] == sorted(owner.values())
    assert sets["bord"][0] == sorted({int(cells[e][p]) for e, k in owner.values()
                                      for p in epos[k]})
    assert sets["coin"][0] == groups[3][2] and nen in (4, 10)


def _mutations(text, rng, count):
    """Single-line corruptions: drop a line, repeat a line, swap two adjacent lines, or replace
    one token with another value."""
    lines = text.split("\n")
    tokens = ["0", "-1", "1e400", "nan", "x", "2.5", "99999", "+3", ".5", "1.", "$Nodes",
              "$EndNodes", '"q"', "4", "11", "15", "8", "-0", "18446744073709551616"]
    for _ in range(count):
        ls = list(lines)
        i = int(rng.integers(len(ls)))
        op = int(rng.integers(4))
        if op == 0:
            del ls[i]
        elif op == 1:
            ls.insert(i, ls[i])
        elif op == 2 and i + 1 < len(ls):
            ls[i], ls[i + 1] = ls[i + 1], ls[i]
        else:
            t = ls[i].split(" ")
            t[int(rng.integers(len(t)))] = tokens[int(rng.integers(len(tokens)))]
            ls[i] = " ".join(t)
        yield "\n".join(ls)


@pytest.mark.parametrize("base", ["triangle", "tet10_line", "generated"])
def test_corrupted_files_get_the_same_answer_from_both_readers(base):
    """Refusal parity: on 300 corrupted copies of a file, both readers accept with identical
    results or refuse with the same variant at the same line."""
    text = {"triangle": cases.TRIANGLE,
            "tet10_line": cases.tet10_file_with_line("1 4 8"),
            "generated": _box_case(True, seed=3, sparse_tags=False)[3]}[base]
    rng = np.random.default_rng(11)
    kinds = set()
    for t in _mutations(text, rng, 300):
        a, b = outcome("library", t), outcome("oracle", t)
        assert same(a, b), (a[:3], b[:3])
        kinds.add(a[1] if a[0] == "error" else "ok")
    assert len(kinds) >= 3


def test_empty_and_whitespace_inputs():
    """An empty text misses $MeshFormat; CRLF line ends and blank lines read as LF."""
    for reader in READERS:
        assert outcome(reader, "")[:2] == ("error", "missing_section")
    crlf = cases.tet10_file().replace("\n", "\r\n\n")
    a, b = outcome("library", crlf), outcome("oracle", crlf)
    assert same(a, b) and a[0] == "ok"
    assert same(a, ("ok",) + _norm_library(lm.read(cases.tet10_file())))
