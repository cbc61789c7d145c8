"""The refusal contract of the library's model creation, on the CPU.

ml_model_create validates the mesh and the materials (Mesh::new, mesh.rs:257-348;
Material::new and with_poisson, material.rs:57-60,145-153) and picks the solver
(SolverKind::Auto, solve.rs:185-194) before any device work, so every refusal here is decided
on the host and these tests run without a GPU. The cases follow the reference's own
tests/solid.rs:282-309 (bad tetrahedra refused with their index).
"""
import numpy as np
import pytest

import maille_amd as ma
from maille_amd import meshes

# tests/solid.rs:284-290
NODES = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0],
                  [1.0, 1.0, 0.0]])


def _problem(nodes, conn, young=210000.0, nu=0.3):
    nodes = np.asarray(nodes, dtype=np.float64)
    conn = np.asarray(conn, dtype=np.int64)
    ne, ndof = len(conn), 3 * len(nodes)
    fixed = np.zeros(ndof, dtype=bool)
    fixed[:3] = True
    return meshes.Problem("refusal", nodes, conn, np.broadcast_to(young, (ne,)).astype(float),
                          np.broadcast_to(nu, (ne,)).astype(float), fixed, np.zeros(ndof),
                          np.zeros(ndof))


def _refusal(problem):
    with pytest.raises(ma.InvalidMesh) as e:
        ma.Model(problem)
    return ma.InvalidMesh.KINDS[e.value.kind], e.value.element


def test_inverted_tetrahedron_is_refused_with_its_index():
    """solid.rs:292-295: InvertedTetrahedron { element: 1 }."""
    assert _refusal(_problem(NODES, [[0, 1, 2, 3], [0, 2, 1, 3]])) == ("inverted tetrahedron", 1)


def test_degenerate_tetrahedron_is_refused_with_its_index():
    """solid.rs:296-299: DegenerateTetrahedron { element: 0 } (four points in z = 0)."""
    assert _refusal(_problem(NODES, [[0, 1, 2, 4]])) == ("degenerate tetrahedron", 0)


def test_curved_tet10_edge_is_refused_with_its_index():
    """solid.rs:301-308: the mid-node of edge 0-1 moved by 0.1 in y gives
    CurvedTetrahedron { element: 0 }."""
    nodes, conn = meshes.to_quadratic_tet(NODES[:4], np.array([[0, 1, 2, 3]]))
    nodes = nodes.copy()
    nodes[conn[0, 4], 1] += 0.1  # mid-node 4 sits on edge 0-1 (shape.rs:135)
    assert _refusal(_problem(nodes, conn)) == ("curved Tet10 edge", 0)
    # unmoved, the same element is accepted by the checks (the refusal is the moved node)
    ok_nodes, _ = meshes.to_quadratic_tet(NODES[:4], np.array([[0, 1, 2, 3]]))
    assert np.abs(ok_nodes[conn[0, 4]] - (NODES[0] + NODES[1]) / 2).max() == 0.0


def test_non_finite_coordinate_is_refused_with_the_node():
    """mesh.rs:261-266: NonFiniteCoordinate { node } (the first such node)."""
    nodes = NODES.copy()
    nodes[3, 2] = np.nan
    nodes[4, 0] = np.inf
    assert _refusal(_problem(nodes, [[0, 1, 2, 4], [0, 1, 2, 3]])) == ("non-finite coordinate", 3)


def test_unknown_node_is_refused_with_the_node():
    """mesh.rs:274-278: UnknownNode { node } (the first out-of-range node of the first element
    that has one)."""
    assert _refusal(_problem(NODES, [[0, 1, 2, 3], [0, 1, 7, 9]])) == ("unknown node", 7)


@pytest.mark.parametrize("young,nu", [(0.0, 0.3), (-5.0, 0.3), (np.nan, 0.3), (1e5, 0.5),
                                      (1e5, -1.0), (1e5, np.nan)])
def test_bad_material_is_refused_with_its_element(young, nu):
    """Material::new (E > 0 and finite, material.rs:57-60) and with_poisson (-1 < nu < 0.5,
    material.rs:145-153); the library reports the first element with a bad material."""
    p = _problem(NODES, [[0, 1, 2, 3], [0, 1, 2, 3]])
    p.young = np.array([210000.0, young])
    p.nu = np.array([0.3, nu])
    assert _refusal(p) == ("bad material", 1)


def test_empty_mesh_is_refused():
    """mesh.rs:258-260: EmptyMesh."""
    p = _problem(NODES, np.zeros((0, 4), dtype=np.int64))
    assert _refusal(p)[0] == "empty mesh"


def test_auto_below_iterative_from_is_answered_as_direct():
    """SolverKind::Auto is Direct below ITERATIVE_FROM = 100,000 free DOFs (solve.rs:185-194);
    the library is the iterative arm, so it answers ML_DIRECT_REQUIRED before any device work
    and the caller takes its skyline arm (tests/iterative.rs:86-92 expects Auto == Direct)."""
    p = meshes.cantilever_tet4()
    assert p.nfree < ma.ITERATIVE_FROM
    with pytest.raises(ma.DirectRequired, match="Direct"):
        ma.Model(p, solver=ma.AUTO)


def test_direct_is_refused():
    p = meshes.cantilever_tet4()
    with pytest.raises(ValueError, match="skyline"):
        ma.Model(p, solver=ma.DIRECT)


def test_auto_from_iterative_from_reaches_the_device_path():
    """At ITERATIVE_FROM free DOFs and above, Auto is the iterative solver: creation goes past
    the solver check (here, without a GPU, it then fails on the device, not on the choice)."""
    p = meshes.jittered_cube(24)  # 45,000 free DOFs: still Direct
    with pytest.raises(ma.DirectRequired):
        ma.Model(p, solver=ma.AUTO)
    big = meshes.cantilever_tet10(n=(40, 8, 8), size=(200.0, 40.0, 40.0))
    assert big.nfree < ma.ITERATIVE_FROM
    q = meshes.jittered_cube(33)  # 3 * 34 * 34 * 33 = 114,444 free DOFs
    assert q.nfree >= ma.ITERATIVE_FROM
    if ma.device_count() > 0:
        pytest.skip("with a device the model is created; covered by the GPU tests")
    with pytest.raises(ma.MailleError) as e:
        ma.Model(q, solver=ma.AUTO)
    assert not isinstance(e.value, ma.DirectRequired)
