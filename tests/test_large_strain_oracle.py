"""Pins the oracle's restatement of the large-strain Newton path (Kinematics::Large:
Saint-Venant-Kirchhoff and Neo-Hookean total-Lagrangian tetrahedra, follower pressures) without
running the reference: the reference's own finite-strain tests restated for Tet4 and Tet10
(tests/finite_strain.rs), plus finite-difference checks of the element tangent and of the
follower-pressure stiffness, which a dropped geometric term or a wrong sign fails."""
import numpy as np
import pytest

import oracle
from maille_amd import meshes as M

E, NU, L, W, D = 210e3, 0.3, 2.0, 1.0, 1.0
LAMBDA = E * NU / ((1 + NU) * (1 - 2 * NU))
MU = E / (2 * (1 + NU))


def solid_meshes():
    """tests/finite_strain.rs:43-59: box_grid([3, 2, 2], [L, W, D]) in Tet4 and Tet10."""
    nodes, tets = M.box_grid((3, 2, 2), (L, W, D))
    qn, qt = M.to_quadratic_tet(nodes, tets)
    return [("Tet4", nodes, tets), ("Tet10", qn, qt)]


def near(a, b):
    return abs(a - b) < 1e-9 * L


def imposed_map(nodes, fn):
    """tests/finite_strain.rs:109-120: every node follows x -> fn(x)."""
    fixed = np.ones(3 * len(nodes), dtype=np.uint8)
    imposed = np.concatenate([np.asarray(fn(n)) - n for n in nodes])
    return fixed, imposed


def free_lateral_tension(nodes, delta):
    """tests/finite_strain.rs:186-204."""
    fixed = np.zeros(3 * len(nodes), dtype=np.uint8)
    imposed = np.zeros(3 * len(nodes))
    for i, n in enumerate(nodes):
        if near(n[0], 0.0):
            fixed[3 * i] = 1
        elif near(n[0], L):
            fixed[3 * i] = 1
            imposed[3 * i] = delta
        if near(n[1], 0.0):
            fixed[3 * i + 1] = 1
        if near(n[2], 0.0):
            fixed[3 * i + 2] = 1
    return fixed, imposed


def run(nodes, conn, fixed, imposed, increments, neo=False, pressures=(), iterative=False):
    return oracle.large_strain(nodes, conn, E, NU, fixed, imposed, np.zeros(3 * len(nodes)), 1.0,
                               increments, [(0.0, 0.0), (1.0, 1.0)], neo_hookean=neo,
                               pressures=pressures, iterative=iterative)


def svk_uniaxial(stretch):
    """tests/finite_strain.rs:147-164 (solid)."""
    exx = (stretch * stretch - 1.0) / 2.0
    eyy = -LAMBDA * exx / (2.0 * (LAMBDA + MU))
    lateral = np.sqrt(1.0 + 2.0 * eyy)
    sxx = LAMBDA * (exx + 2.0 * eyy) + 2.0 * MU * exx
    return lateral, stretch * stretch * sxx / (stretch * lateral * lateral)


def neo_principal(stretch, j):
    return (MU * (stretch * stretch - 1.0) + LAMBDA * np.log(j)) / j


def neo_uniaxial(stretch):
    """tests/finite_strain.rs:375-392 (bisection of the lateral stretch)."""
    lo, hi = 0.01, 10.0
    f = lambda l: neo_principal(l, stretch * l * l)  # noqa: E731
    sign = np.sign(f(lo))
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.sign(f(mid)) == sign:
            lo = mid
        else:
            hi = mid
    lat = 0.5 * (lo + hi)
    return lat, neo_principal(stretch, stretch * lat * lat)


def rotation(n):
    x, y, z = -n[1], n[0], n[2]
    return [x, -z, y]


@pytest.mark.parametrize("neo", [False, True], ids=["SVK", "NeoHookean"])
@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_a_rigid_rotation_leaves_no_stress(mesh, neo):
    """tests/finite_strain.rs:122-145 and :480-496."""
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = imposed_map(nodes, rotation)
    r = run(nodes, conn, fixed, imposed, 1, neo)
    assert r["status"] == oracle.OK
    assert np.abs(r["stresses"][1][..., 3:]).max() < 1e-10 * E
    assert np.abs(r["reactions"][1]).max() < 1e-10 * E


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_a_prescribed_uniaxial_stretch_gives_the_closed_form_stress(mesh):
    """tests/finite_strain.rs:166-185 (SVK, stretch 1.4)."""
    name, nodes, conn = solid_meshes()[mesh]
    lat, sxx = svk_uniaxial(1.4)
    fixed, imposed = imposed_map(nodes, lambda n: [1.4 * n[0], lat * n[1], lat * n[2]])
    s = run(nodes, conn, fixed, imposed, 1)["stresses"][1]
    assert np.abs(s[..., 3] - sxx).max() < 1e-9 * sxx
    assert np.abs(s[..., 4:]).max() < 1e-9 * sxx


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_newton_finds_the_uniaxial_closed_form_with_free_lateral_sides(mesh):
    """tests/finite_strain.rs:206-219 (SVK, stretch 1.5, 5 increments)."""
    name, nodes, conn = solid_meshes()[mesh]
    _, sxx = svk_uniaxial(1.5)
    fixed, imposed = free_lateral_tension(nodes, 0.5 * L)
    s = run(nodes, conn, fixed, imposed, 5)["stresses"][5]
    assert np.abs(s[..., 3] - sxx).max() < 1e-6 * sxx
    assert np.abs(s[..., 4]).max() < 1e-6 * sxx


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_newton_converges_quadratically(mesh):
    """tests/finite_strain.rs:221-247, with PCG and skyline taking the same Newton path."""
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = free_lateral_tension(nodes, 0.5 * L)
    d = run(nodes, conn, fixed, imposed, 3)
    i = run(nodes, conn, fixed, imposed, 3, iterative=True)
    longest = 0
    for rd, ri in zip(d["residuals"], i["residuals"]):
        assert len(rd) <= 8 and len(rd) == len(ri)
        longest = max(longest, len(rd))
        for a, b in zip(rd[1:], rd[2:]):
            if a < 1e-2:
                assert b <= max(10 * a * a, 1e-12)
    assert longest >= 4
    assert np.abs(i["u"][-1] - d["u"][-1]).max() < 1e-8 * L


@pytest.mark.parametrize("neo", [False, True], ids=["SVK", "NeoHookean"])
@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_a_small_load_gives_the_linear_solution(mesh, neo):
    """tests/finite_strain.rs:249-277 and :498-523."""
    name, nodes, conn = solid_meshes()[mesh]
    eps = 1e-6
    fixed, imposed = free_lateral_tension(nodes, eps * L)
    prob = M.Problem(name, nodes, conn, np.full(len(conn), E), np.full(len(conn), NU),
                     fixed.astype(bool), imposed, np.zeros(3 * len(nodes)))
    lin = oracle.model_solve(prob, iterative=False)
    r = run(nodes, conn, fixed, imposed, 1, neo)
    assert np.abs(r["u"][1] - lin["displacements"]).max() < 1e-5 * eps * L
    assert np.abs(r["stresses"][1][..., 3:5] - lin["stresses"][..., 3:5]).max() < 1e-5 * E * eps


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_neo_hookean_simple_shear_and_uniaxial_tension_match_the_closed_forms(mesh):
    """tests/finite_strain.rs:433-478."""
    name, nodes, conn = solid_meshes()[mesh]
    g = 0.8
    fixed, imposed = imposed_map(nodes, lambda n: [n[0] + g * n[1], n[1], n[2]])
    s = run(nodes, conn, fixed, imposed, 1, True)["stresses"][1]
    assert np.abs(s[..., 3] - MU * g * g).max() < 1e-9 * MU
    assert np.abs(s[..., 8] - MU * g).max() < 1e-9 * MU
    assert np.abs(s[..., [4, 5, 6, 7]]).max() < 1e-9 * MU
    lat, sxx = neo_uniaxial(2.0)
    fixed, imposed = imposed_map(nodes, lambda n: [2.0 * n[0], lat * n[1], lat * n[2]])
    s = run(nodes, conn, fixed, imposed, 4, True)["stresses"][4]
    assert np.abs(s[..., 3] - sxx).max() < 1e-8 * sxx and np.abs(s[..., 4]).max() < 1e-8 * sxx
    fixed, imposed = free_lateral_tension(nodes, 1.0 * L)
    s = run(nodes, conn, fixed, imposed, 8, True)["stresses"][8]
    assert np.abs(s[..., 3] - sxx).max() < 1e-6 * sxx and np.abs(s[..., 4]).max() < 1e-6 * sxx


def hydrostatic(nodes, conn, p):
    """tests/finite_strain.rs:543-577: symmetry supports, follower pressure on x = L, y = W,
    z = D."""
    fixed = np.zeros(3 * len(nodes), dtype=np.uint8)
    for i, n in enumerate(nodes):
        for c in range(3):
            if near(n[c], 0.0):
                fixed[3 * i + c] = 1
    faces = ((0, 1, 2), (0, 3, 1), (1, 3, 2), (2, 3, 0))
    pressures = []
    for e, t in enumerate(conn):
        for fc, vs in enumerate(faces):
            corners = nodes[t[list(vs)]]
            if (all(near(c[0], L) for c in corners) or all(near(c[1], W) for c in corners)
                    or all(near(c[2], D) for c in corners)):
                pressures.append((e, fc, p))
    return fixed, pressures


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_a_following_pressure_compresses_a_neo_hookean_homogeneously(mesh):
    """tests/finite_strain.rs:579-612: sigma = -p at a homogeneous stretch of 0.8."""
    name, nodes, conn = solid_meshes()[mesh]
    stretch = 0.8
    p = -neo_principal(stretch, stretch ** 3)
    fixed, pressures = hydrostatic(nodes, conn, p)
    r = run(nodes, conn, fixed, np.zeros(3 * len(nodes)), 4, True, pressures)
    assert r["status"] == oracle.OK
    u = r["u"][4].reshape(-1, 3)
    assert np.abs(u - (stretch - 1.0) * nodes).max() < 1e-6 * L
    s = r["stresses"][4]
    assert np.abs(s[..., 3] + p).max() < 1e-6 * p and np.abs(s[..., 4] + p).max() < 1e-6 * p
    assert np.abs(s[..., 8]).max() < 1e-6 * p


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_newton_cuts_the_increment_and_a_collapse_is_a_typed_error(mesh):
    """tests/finite_strain.rs:614-650: a stretch of 4 in one increment is reached by cutting; a
    compression to -0.2 L ends on NotConverged or Singular."""
    name, nodes, conn = solid_meshes()[mesh]
    _, sxx = neo_uniaxial(4.0)
    fixed, imposed = free_lateral_tension(nodes, 3.0 * L)
    r = run(nodes, conn, fixed, imposed, 1, True)
    assert r["status"] == oracle.OK and r["cuts"] > 0
    assert np.abs(r["stresses"][1][..., 3] - sxx).max() < 1e-6 * sxx
    fixed, imposed = free_lateral_tension(nodes, -1.2 * L)
    r = run(nodes, conn, fixed, imposed, 1, True)
    assert r["status"] in (oracle.NEWTON_NOT_CONVERGED, oracle.SINGULAR)


@pytest.mark.parametrize("neo", [False, True], ids=["SVK", "NeoHookean"])
@pytest.mark.parametrize("nen", [4, 10])
def test_element_tangent_is_the_derivative_of_the_internal_force(nen, neo):
    """finite_strain.rs:119-214: tangent = d fint / d u_e (central differences) at a large
    deformation; a dropped geometric term fails this."""
    rng = np.random.default_rng(11)
    x4 = np.array([[0, 0, 0], [1.5, 0.1, 0], [0.2, 1.2, 0.1], [0.1, 0.2, 1.1]]) + rng.normal(0, .03, (4, 3))
    x = x4 if nen == 4 else np.vstack([x4, [(x4[a] + x4[b]) / 2 for a, b in M.TET10_EDGES]])
    ue = rng.normal(0, 0.1, 3 * nen)
    fint, tan, _ = oracle.large_response(x, E, NU, ue, neo)
    h = 1e-7
    for c in range(3 * nen):
        dp, dm = ue.copy(), ue.copy()
        dp[c] += h
        dm[c] -= h
        num = (oracle.large_response(x, E, NU, dp, neo)[0]
               - oracle.large_response(x, E, NU, dm, neo)[0]) / (2 * h)
        assert np.abs(tan[:, c] - num).max() < 1e-5 * np.abs(tan).max()


@pytest.mark.parametrize("nen", [4, 10])
def test_follower_stiffness_is_the_derivative_of_the_follower_forces(nen):
    """assembly.rs:732-862: stiffness = d forces / d u (central differences) on a deformed face,
    and the forces of an undeformed face sum to -p A n."""
    nodes, conn = M.box_grid((1, 1, 1), (1.0, 1.0, 1.0))
    if nen == 10:
        nodes, conn = M.to_quadratic_tet(nodes, conn)
    rng = np.random.default_rng(5)
    e, face, p = 0, 0, 3.0
    pos, f0, _ = oracle.follower_pressure(nodes, conn[e], face, p, np.zeros(3 * len(nodes)))
    corners = nodes[conn[e][[0, 1, 2]]]
    cross = np.cross(corners[1] - corners[0], corners[2] - corners[0])
    outward = corners.mean(axis=0) - nodes[conn[e][:4]].mean(axis=0)
    n = cross / np.linalg.norm(cross) * np.sign(cross @ outward)
    area = np.linalg.norm(cross) / 2
    assert np.abs(f0.reshape(-1, 3).sum(axis=0) + p * area * n).max() < 1e-12 * p * area
    u = rng.normal(0, 0.05, 3 * len(nodes))
    _, f, k = oracle.follower_pressure(nodes, conn[e], face, p, u)
    h = 1e-7
    for b in range(len(pos)):
        for c in range(3):
            dof = 3 * conn[e][pos[b]] + c
            up, um = u.copy(), u.copy()
            up[dof] += h
            um[dof] -= h
            num = (oracle.follower_pressure(nodes, conn[e], face, p, up)[1]
                   - oracle.follower_pressure(nodes, conn[e], face, p, um)[1]) / (2 * h)
            assert np.abs(k[:, 3 * b + c] - num).max() < 1e-6 * max(np.abs(k).max(), 1.0)
