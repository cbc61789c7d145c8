"""Device parity of the large-strain Newton path (Kinematics::Large: Saint-Venant-Kirchhoff and
Neo-Hookean tetrahedra, follower pressures) against the oracle's restatement, on the reference's
finite-strain scenarios (tests/finite_strain.rs). The element responses follow the oracle's
operations in order, the follower terms are host code restating the same function, and the PCG
dot products and residual norms are tree sums: displacements within 1e-10 (relative L2) of the
oracle's PCG-driven Newton for Saint-Venant-Kirchhoff, 1e-9 for Neo-Hookean (whose ln J is the
device's log against the host's, which may differ in the last bit), and the same Newton
iteration counts."""
import numpy as np
import pytest

import oracle
from maille_amd import meshes as M
from maille_amd import plasticity as P
from conftest import rel_l2
from test_large_strain_oracle import (E, L, NU, free_lateral_tension, hydrostatic, neo_principal,
                                      neo_uniaxial, solid_meshes, svk_uniaxial)

pytestmark = pytest.mark.gpu


def device(gpu, nodes, conn, fixed, imposed, increments, neo=False, pressures=()):
    prob = M.Problem("large", nodes, conn, np.full(len(conn), E), np.full(len(conn), NU),
                     np.asarray(fixed, dtype=bool), imposed, np.zeros(3 * len(nodes)))
    model = gpu.Model(prob)
    return P.solve_large_strain(model, P.Step(1.0, increments), P.Amplitude.ramp(1.0),
                                neo_hookean=neo, pressures=pressures)


def reference(nodes, conn, fixed, imposed, increments, neo=False, pressures=()):
    return oracle.large_strain(nodes, conn, E, NU, fixed, imposed, np.zeros(3 * len(nodes)), 1.0,
                               increments, [(0.0, 0.0), (1.0, 1.0)], neo_hookean=neo,
                               pressures=pressures, iterative=True)


def compare(run, incs, info, tol):
    assert run["status"] == oracle.OK
    assert info["newton_iterations"] == run["newton"] and info["cuts"] == run["cuts"]
    for k, inc in enumerate(incs):
        if k:
            assert len(inc.residuals) == len(run["residuals"][k - 1])
            np.testing.assert_allclose(inc.residuals, run["residuals"][k - 1], rtol=1e-5,
                                       atol=1e-11)
        assert rel_l2(inc.displacements, run["u"][k]) <= tol, k
        scale = max(np.abs(run["stresses"][k][..., 3:]).max(), 1.0)
        assert np.abs(inc.stresses - run["stresses"][k]).max() <= 1e-7 * scale, k


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_svk_tension_with_free_lateral_sides_matches_the_oracle(gpu, mesh):
    """tests/finite_strain.rs:206-247: Newton to the closed form over 5 increments, converging
    quadratically, step by step as the oracle."""
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = free_lateral_tension(nodes, 0.5 * L)
    incs, info = device(gpu, nodes, conn, fixed, imposed, 5)
    run = reference(nodes, conn, fixed, imposed, 5)
    compare(run, incs, info, 1e-10)
    _, sxx = svk_uniaxial(1.5)
    assert np.abs(incs[5].stresses[..., 3] - sxx).max() < 1e-6 * sxx
    assert max(len(i.residuals) for i in incs[1:]) >= 3


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_neo_hookean_follower_pressure_compression_matches_the_oracle(gpu, mesh):
    """tests/finite_strain.rs:579-612: follower pressures on three faces compress a Neo-Hookean
    block homogeneously to a stretch of 0.8, sigma = -p."""
    name, nodes, conn = solid_meshes()[mesh]
    stretch = 0.8
    p = -neo_principal(stretch, stretch ** 3)
    fixed, pressures = hydrostatic(nodes, conn, p)
    zero = np.zeros(3 * len(nodes))
    incs, info = device(gpu, nodes, conn, fixed, zero, 4, True, pressures)
    run = reference(nodes, conn, fixed, zero, 4, True, pressures)
    compare(run, incs, info, 1e-9)
    u = incs[4].displacements.reshape(-1, 3)
    assert np.abs(u - (stretch - 1.0) * nodes).max() < 1e-6 * L
    s = incs[4].stresses
    assert np.abs(s[..., 3] + p).max() < 1e-6 * p and np.abs(s[..., 5] + p).max() < 1e-6 * p


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_neo_hookean_stretch_of_four_is_reached_by_cutting(gpu, mesh):
    """tests/finite_strain.rs:614-630: one increment too large for Newton from rest, saved by the
    cuts, as in the oracle."""
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = free_lateral_tension(nodes, 3.0 * L)
    incs, info = device(gpu, nodes, conn, fixed, imposed, 1, True)
    run = reference(nodes, conn, fixed, imposed, 1, True)
    assert info["cuts"] > 0
    compare(run, incs, info, 1e-9)
    _, sxx = neo_uniaxial(4.0)
    assert np.abs(incs[1].stresses[..., 3] - sxx).max() < 1e-6 * sxx


def test_a_small_load_gives_the_linear_solution_on_the_device(gpu):
    """tests/finite_strain.rs:249-277 on the device (Tet10, Saint-Venant-Kirchhoff)."""
    name, nodes, conn = solid_meshes()[1]
    eps = 1e-6
    fixed, imposed = free_lateral_tension(nodes, eps * L)
    prob = M.Problem(name, nodes, conn, np.full(len(conn), E), np.full(len(conn), NU),
                     fixed.astype(bool), imposed, np.zeros(3 * len(nodes)))
    model = gpu.Model(prob)
    u_lin, _, _ = model.solve()
    incs, _ = P.solve_large_strain(model, P.Step(1.0, 1), P.Amplitude.ramp(1.0))
    assert np.abs(incs[1].displacements - u_lin).max() < 1e-5 * eps * L
    u_again, _, _ = model.solve()
    assert rel_l2(u_again, u_lin) <= 1e-12


@pytest.mark.parametrize("neo", [False, True], ids=["SVK", "NeoHookean"])
@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_every_node_imposed_rigid_rotation_and_stretch(gpu, mesh, neo):
    """tests/finite_strain.rs:122-185 and :480-496 on the device: every DOF imposed (no free
    system, as FreeSystem::new allows); a rigid rotation leaves no stress and no reaction, a
    prescribed stretch gives the oracle's stresses."""
    from test_large_strain_oracle import imposed_map, rotation
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = imposed_map(nodes, rotation)
    incs, info = device(gpu, nodes, conn, fixed, imposed, 1, neo)
    assert np.abs(incs[1].stresses[..., 3:]).max() < 1e-10 * E
    assert np.abs(incs[1].reactions).max() < 1e-10 * E
    lat, sxx = svk_uniaxial(1.4)
    fixed, imposed = imposed_map(nodes, lambda n: [1.4 * n[0], lat * n[1], lat * n[2]])
    incs, info = device(gpu, nodes, conn, fixed, imposed, 1, neo)
    run = reference(nodes, conn, fixed, imposed, 1, neo)
    assert np.array_equal(incs[1].displacements, run["u"][1])
    assert np.abs(incs[1].stresses - run["stresses"][1]).max() <= 1e-9 * abs(sxx)
    assert len(incs[1].residuals) == len(run["residuals"][0]) == 2
    if not neo:
        assert np.abs(incs[1].stresses[..., 3] - sxx).max() < 1e-9 * sxx


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_neo_hookean_follower_compression_with_jacobi_sweeps(gpu, mesh):
    """The large-strain Newton with the FP32 factor applied by Jacobi sweeps (ML_MIXED_JACOBI):
    every PCG solve meets the 1e-12 residual, so Newton takes the oracle's iterations and ends
    within 1e-8 (relative L2) of the oracle's FP64 run and of the homogeneous closed form."""
    name, nodes, conn = solid_meshes()[mesh]
    stretch = 0.8
    p = -neo_principal(stretch, stretch ** 3)
    fixed, pressures = hydrostatic(nodes, conn, p)
    zero = np.zeros(3 * len(nodes))
    prob = M.Problem("large", nodes, conn, np.full(len(conn), E), np.full(len(conn), NU),
                     np.asarray(fixed, dtype=bool), zero, zero.copy())
    model = gpu.Model(prob, precision=gpu.MIXED_JACOBI)
    incs, info = P.solve_large_strain(model, P.Step(1.0, 4), P.Amplitude.ramp(1.0),
                                      neo_hookean=True, pressures=pressures)
    run = reference(nodes, conn, fixed, zero, 4, True, pressures)
    assert run["status"] == oracle.OK
    assert info["newton_iterations"] == run["newton"] and info["cuts"] == run["cuts"]
    for k, inc in enumerate(incs):
        assert rel_l2(inc.displacements, run["u"][k]) <= 1e-8, k
    u = incs[4].displacements.reshape(-1, 3)
    assert np.abs(u - (stretch - 1.0) * nodes).max() < 1e-6 * L
