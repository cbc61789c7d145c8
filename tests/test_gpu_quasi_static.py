"""Device parity of the quasi-static elasto-plastic path (Model::solve_quasi_static, model.rs:
715-1022) against the oracle's restatement, on the reference's own plasticity scenarios
(tests/plasticity.rs) and a bending bar with a growing plastic zone. The element responses are
the oracle's operation for operation; the PCG dot products and the residual norms are tree
sums, so Newton iterates agree to round-off: displacements within 1e-10 (relative L2) of the
oracle's PCG-driven Newton, the same Newton iteration counts, relative residuals within 1e-6
(relative) above the 1e-11 round-off floor of a converged state, p within 1e-12."""
import numpy as np
import pytest

import oracle
from maille_amd import meshes as M
from maille_amd import plasticity as P
from conftest import rel_l2
from test_plasticity_oracle import (D, E, H, L, NU, SOLID_STRAIN, SY, W, boundary_history, near,
                                    solid_meshes, tension_bcs, uniaxial_closed_form)

pytestmark = pytest.mark.gpu


def model_of(gpu, name, nodes, conn, fixed, imposed, f=None):
    prob = M.Problem(name, nodes, conn, np.full(len(conn), E), np.full(len(conn), NU),
                     np.asarray(fixed, dtype=bool), imposed,
                     np.zeros(3 * len(nodes)) if f is None else f)
    return gpu.Model(prob), prob


def compare(run, incs, tol=1e-10):
    assert run["status"] == oracle.OK
    for k, inc in enumerate(incs):
        if k:
            assert len(inc.residuals) == len(run["residuals"][k - 1]), k
            # residuals above round-off agree; those at round-off (a converged state) differ
            np.testing.assert_allclose(inc.residuals, run["residuals"][k - 1],
                                       rtol=1e-6, atol=1e-11)
        assert rel_l2(inc.displacements, run["u"][k]) <= tol, k
        assert np.abs(inc.equivalent_plastic_strains - run["p"][k]).max() <= 1e-12, k
        scale = max(np.abs(run["stresses"][k][..., 3:]).max(), 1.0)
        assert np.abs(inc.stresses[..., 3:] - run["stresses"][k][..., 3:]).max() <= 1e-8 * scale
        assert np.array_equal(inc.stresses[..., :3], run["stresses"][k][..., :3])
        assert np.abs(inc.reactions - run["reactions"][k]).max() <= 1e-8 * max(
            np.abs(run["reactions"][k]).max(), 1.0)


@pytest.mark.parametrize("hardening", [H, 0.0])
@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_uniaxial_tension_loads_and_unloads_on_the_device(gpu, mesh, hardening):
    """tests/plasticity.rs:157-228 on the device, against the oracle and the closed form."""
    name, nodes, conn = solid_meshes()[mesh]
    strain_max = 3.0 * SY / E
    sigma_max, p_max = uniaxial_closed_form(strain_max, hardening)
    last = p_max / strain_max
    fixed, imposed = tension_bcs(nodes, strain_max * L)
    amp = [(0.0, 0.0), (1.0, 1.0), (2.0, last)]
    model, _ = model_of(gpu, name, nodes, conn, fixed, imposed)
    incs, info = P.solve_quasi_static(model, P.Step(2.0, 8), P.Amplitude(amp), SY, hardening)
    run = oracle.quasi_static(nodes, conn, E, NU, SY, hardening, fixed, imposed,
                              np.zeros(3 * len(nodes)), 2.0, 8, amp, iterative=True)
    compare(run, incs)
    end = [i for i, n in enumerate(nodes) if near(n[0], L)]
    for k, inc in enumerate(incs):
        t = 0.25 * k
        assert inc.time == t
        lam = t if t <= 1.0 else 1.0 + (last - 1.0) * (t - 1.0)
        strain = lam * strain_max
        sigma, p = (uniaxial_closed_form(strain, hardening) if t <= 1.0
                    else (sigma_max - E * (strain_max - strain), p_max))
        force = sum(inc.reactions[3 * i] for i in end)
        assert abs(force - sigma * W * D) < 1e-6 * SY
        assert np.abs(inc.equivalent_plastic_strains - p).max() < 1e-9
    assert info["newton_iterations"] == run["newton"] and info["cuts"] == 0


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_homogeneous_strain_history_on_the_device(gpu, mesh):
    """tests/plasticity.rs:239-321: load, then unload beyond zero (reverse yielding)."""
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = boundary_history(nodes, SOLID_STRAIN)
    amp = [(0.0, 0.0), (1.0, 1.0), (2.0, -0.5)]
    model, _ = model_of(gpu, name, nodes, conn, fixed, imposed)
    incs, _ = P.solve_quasi_static(model, P.Step(2.0, 8), P.Amplitude(amp), SY, H)
    run = oracle.quasi_static(nodes, conn, E, NU, SY, H, fixed, imposed, np.zeros(3 * len(nodes)),
                              2.0, 8, amp, iterative=True)
    compare(run, incs)
    state = np.zeros(7)
    for k in range(1, 9):
        t = 0.25 * k
        lam = t if t <= 1.0 else 1.0 - 1.5 * (t - 1.0)
        s, state, _, _ = oracle.von_mises(E, NU, SY, H, np.array(SOLID_STRAIN) * lam, state)
        assert np.abs(incs[k].stresses[..., 3:] - s).max() < 1e-6 * SY
    assert state[6] > 0


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_elastic_model_matches_the_linear_solve_in_one_iteration(gpu, mesh):
    """tests/plasticity.rs:322-370 on the device; afterwards the model's linear solve still
    gives the elastic answer (K is reassembled after the tangents)."""
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = tension_bcs(nodes, 0.0)
    for i, n in enumerate(nodes):
        if near(n[0], L):
            fixed[3 * i] = 0
    f = np.zeros(3 * len(nodes))
    for i, n in enumerate(nodes):
        if near(n[0], L) and near(n[1], W):
            f[3 * i] = 10.0
    model, prob = model_of(gpu, name, nodes, conn, fixed, imposed, f)
    u_lin, r_lin, _ = model.solve()
    incs, _ = P.solve_quasi_static(model, P.Step(1.0, 2), P.Amplitude.ramp(1.0), np.inf, 0.0)
    for k in (1, 2):
        assert len(incs[k].residuals) == 2
        assert np.abs(incs[k].displacements - 0.5 * k * u_lin).max() < 1e-9 * 1e-3
    assert np.abs(incs[2].reactions - r_lin).max() < 1e-6
    assert np.abs(incs[2].equivalent_plastic_strains).max() < 1e-15
    u_again, _, _ = model.solve()
    assert rel_l2(u_again, u_lin) <= 1e-12


def bending_bar(nen):
    nodes, conn = M.box_grid((6, 2, 2), (2 * L, W, D))
    if nen == 10:
        nodes, conn = M.to_quadratic_tet(nodes, conn)
    fixed = np.zeros(3 * len(nodes), dtype=np.uint8)
    imposed = np.zeros(3 * len(nodes))
    for i, n in enumerate(nodes):
        if near(n[0], 0.0):
            fixed[3 * i:3 * i + 3] = 1
        elif near(n[0], 2 * L):
            fixed[3 * i + 2] = 1
            imposed[3 * i + 2] = 0.05
    return nodes, conn, fixed, imposed


@pytest.mark.parametrize("nen", [4, 10])
def test_bending_with_a_growing_plastic_zone_matches_the_oracle(gpu, nen):
    """tests/plasticity.rs:371-419 (restated on tetrahedra): a clamped bar bent by an imposed
    tip displacement, five increments; Newton converges quadratically and the device follows
    the oracle's Newton iteration by iteration."""
    nodes, conn, fixed, imposed = bending_bar(nen)
    amp = [(0.0, 0.0), (1.0, 1.0)]
    model, _ = model_of(gpu, f"bar{nen}", nodes, conn, fixed, imposed)
    incs, info = P.solve_quasi_static(model, P.Step(1.0, 5), P.Amplitude(amp), SY, H)
    run = oracle.quasi_static(nodes, conn, E, NU, SY, H, fixed, imposed, np.zeros(3 * len(nodes)),
                              1.0, 5, amp, iterative=True)
    compare(run, incs, tol=1e-9)
    assert max(len(i.residuals) for i in incs[1:]) >= 4
    assert incs[-1].equivalent_plastic_strains.max() > 0
    assert info["pcg_iterations"] > 0


def test_a_force_beyond_the_limit_load_fails_on_the_same_increment(gpu):
    """tests/plasticity.rs:420-462 (restated on tetrahedra): a perfectly plastic bar pulled
    above its limit load. The increments before the limit agree with the oracle; the run ends
    on an error in the same increment. Which error it is, and after how many cuts (PlasticCollapse,
    or with PCG a Singular curvature or IterativeNotConverged on the nearly singular tangent,
    which the reference propagates uncut) depends on round-off at the limit, in the reference
    as here, so only its being a MailleError is compared."""
    import maille_amd
    nodes, conn = M.box_grid((4, 1, 1), (L, W, D))
    fixed, imposed = tension_bcs(nodes, 0.0)
    end = [i for i, n in enumerate(nodes) if near(n[0], L)]
    for i in end:
        fixed[3 * i] = 0
    f = np.zeros(3 * len(nodes))
    for i in end:
        f[3 * i] = 1.2 * SY * W * D / len(end)
    run = oracle.quasi_static(nodes, conn, E, NU, SY, 0.0, fixed, imposed, f, 1.0, 6,
                              [(0.0, 0.0), (1.0, 1.0)], iterative=True)
    assert run["status"] != oracle.OK
    model, _ = model_of(gpu, "limit", nodes, conn, fixed, imposed, f)
    with pytest.raises(maille_amd.MailleError) as err:
        P.solve_quasi_static(model, P.Step(1.0, 6), P.Amplitude.ramp(1.0), SY, 0.0)
    assert err.value.increment == run["fail_increment"]
    ok = P.solve_quasi_static(model, P.Step(1.0, 6), P.Amplitude([(0.0, 0.0), (0.6, 0.6),
                                                                    (1.0, 0.6)]), SY, 0.0)[0]
    ref = oracle.quasi_static(nodes, conn, E, NU, SY, 0.0, fixed, imposed, f, 1.0, 6,
                              [(0.0, 0.0), (0.6, 0.6), (1.0, 0.6)], iterative=True)
    compare(ref, ok)


@pytest.mark.parametrize("nen", [4, 10])
def test_bending_with_jacobi_sweeps_follows_the_oracle(gpu, nen):
    """The Newton caller with the FP32 factor applied by Jacobi sweeps (ML_MIXED_JACOBI): each
    PCG solve still meets the reference's 1e-12 residual, so Newton takes the oracle's iterations
    increment by increment, with displacements within 1e-8 (relative L2), plastic strains within
    1e-10 and stresses within 1e-6 (relative) of the oracle's FP64 Newton."""
    nodes, conn, fixed, imposed = bending_bar(nen)
    amp = [(0.0, 0.0), (1.0, 1.0)]
    prob = M.Problem(f"bar{nen}", nodes, conn, np.full(len(conn), E), np.full(len(conn), NU),
                     np.asarray(fixed, dtype=bool), imposed, np.zeros(3 * len(nodes)))
    model = gpu.Model(prob, precision=gpu.MIXED_JACOBI)
    incs, info = P.solve_quasi_static(model, P.Step(1.0, 5), P.Amplitude(amp), SY, H)
    run = oracle.quasi_static(nodes, conn, E, NU, SY, H, fixed, imposed, np.zeros(3 * len(nodes)),
                              1.0, 5, amp, iterative=True)
    assert run["status"] == oracle.OK
    assert info["newton_iterations"] == run["newton"] and info["cuts"] == 0
    for k, inc in enumerate(incs):
        if k:
            assert len(inc.residuals) == len(run["residuals"][k - 1]), k
        assert rel_l2(inc.displacements, run["u"][k]) <= 1e-8, k
        assert np.abs(inc.equivalent_plastic_strains - run["p"][k]).max() <= 1e-10, k
        scale = max(np.abs(run["stresses"][k][..., 3:]).max(), 1.0)
        assert np.abs(inc.stresses[..., 3:] - run["stresses"][k][..., 3:]).max() <= 1e-6 * scale
    assert incs[-1].equivalent_plastic_strains.max() > 0
