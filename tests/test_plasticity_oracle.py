"""Pins the oracle's restatement of the quasi-static elasto-plastic path (von Mises law,
element response, Newton-Raphson with increment cutting) without running the reference: the
reference's own unit and integration tests of that path, restated (plasticity.rs:128-294,
tests/plasticity.rs), plus finite-difference checks of both tangents, which a dropped or
mis-signed term of the consistent tangent fails."""
import numpy as np
import pytest

import oracle
from maille_amd import meshes as M
from maille_amd import plasticity as P

E, NU, SY, H = 210e3, 0.3, 250.0, 2000.0
L, W, D = 2.0, 1.0, 1.0


def law(hardening, yield_stress=SY):
    return lambda strain, prev=None: oracle.von_mises(E, NU, yield_stress, hardening, strain, prev)


def vm(s):
    a, b, c = s[0], s[1], s[2]
    return np.sqrt(a * a + b * b + c * c - a * b - b * c - c * a
                   + 3.0 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2))


def uniaxial_strain(sigma, p):
    lateral = -NU * sigma / E - p / 2.0
    return [sigma / E + p, lateral, lateral, 0.0, 0.0, 0.0]


# ------------------------------------------------------------------ the material point
def test_elastic_tangent_is_solid_d():
    """plasticity.rs:170-181."""
    stress, state, tan, y = law(H)(uniaxial_strain(0.5 * SY, 0.0))
    assert not y
    assert np.abs(tan - oracle.solid_d(E, NU)).max() < 1e-9 * E


def test_uniaxial_tension_follows_the_hardening_slope_then_unloads_elastically():
    """plasticity.rs:183-212."""
    p = 2e-3
    sigma = SY + H * p
    s, state, _, y = law(H)(uniaxial_strain(sigma, p))
    assert y and abs(s[0] - sigma) < 1e-9 * SY and np.abs(s[1:]).max() < 1e-9 * SY
    assert abs(state[6] - p) < 1e-12
    assert abs(state[0] - p) < 1e-12 and abs(state[1] + p / 2) < 1e-12
    assert abs(state[0] + state[1] + state[2]) < 1e-15
    total = uniaxial_strain(sigma, p)
    further, _, _, _ = law(H)(uniaxial_strain(sigma + 10.0, p + 10.0 / H), state)
    de = uniaxial_strain(sigma + 10.0, p + 10.0 / H)[0] - total[0]
    assert abs((further[0] - sigma) - E * H / (E + H) * de) < 1e-9 * SY
    unloaded, ustate, _, uy = law(H)(uniaxial_strain(0.0, p), state)
    assert not uy and np.abs(unloaded).max() < 1e-9 * SY
    assert np.array_equal(ustate, state)


def test_perfect_plasticity_stays_on_the_yield_surface():
    """plasticity.rs:214-229."""
    state = np.zeros(7)
    for k in range(1, 11):
        strain = [2e-3 * k, -1e-3 * k, -0.5e-3 * k, 0.0, 0.0, 3e-3 * k]
        s, state_k, _, y = law(0.0)(strain, state)
        if y:
            assert abs(vm(s) - SY) < 1e-9 * SY
        state = state_k
    assert state[6] > 0


def test_pure_shear_yields_at_sigma_y_over_root_3():
    """plasticity.rs:231-237."""
    g = E / (2 * (1 + NU))
    s, _, _, _ = law(0.0)([0, 0, 0, 0, 0, 10.0 * SY / g])
    assert abs(s[5] - SY / np.sqrt(3.0)) < 1e-9 * SY


def test_pressure_does_not_cause_yielding():
    """plasticity.rs:239-246."""
    e = 1e-2
    s, _, _, y = law(H)([e, e, e, 0, 0, 0])
    bulk = E / (3 * (1 - 2 * NU))
    assert not y and abs(s[0] - 3 * bulk * e) < 1e-9 * bulk * e


def test_consistent_tangent_matches_finite_differences():
    """plasticity.rs:248-281."""
    start = law(H)([1.5e-3, -0.4e-3, -0.2e-3, 0.3e-3, -0.2e-3, 0.8e-3])[1]
    strain = np.array([3e-3, -1e-3, -0.6e-3, 1e-3, -0.5e-3, 2e-3])
    for hardening in (H, 0.0):
        for prev in (np.zeros(7), start):
            s, _, tan, y = law(hardening)(strain, prev)
            assert y
            h = 1e-8
            for j in range(6):
                dp, dm = strain.copy(), strain.copy()
                dp[j] += h
                dm[j] -= h
                num = (law(hardening)(dp, prev)[0] - law(hardening)(dm, prev)[0]) / (2 * h)
                assert np.abs(tan[:, j] - num).max() < 1e-7 * E
            assert np.abs(tan - tan.T).max() < 1e-9 * E


def test_infinite_yield_stress_is_elastic():
    """plasticity.rs:283-293."""
    s, _, _, y = law(H, np.inf)(uniaxial_strain(10.0 * SY, 0.0))
    assert not y and abs(s[0] - 10.0 * SY) < 1e-9 * SY


# ------------------------------------------------------------------ the element response
@pytest.mark.parametrize("nen", [4, 10])
def test_element_response_tangent_is_the_derivative_of_the_internal_force(nen):
    """element.rs:395-452: tangent = d fint / d u_e (finite differences) on a yielding element;
    in the elastic range it is solid_stiffness and fint = Ke u_e."""
    rng = np.random.default_rng(7)
    x4 = np.array([[0, 0, 0], [2, 0.1, 0], [0.3, 1.5, 0.2], [0.1, 0.2, 1.2]]) + rng.normal(0, .05, (4, 3))
    x = x4 if nen == 4 else np.vstack([x4, [(x4[a] + x4[b]) / 2 for a, b in M.TET10_EDGES]])
    m = 3 * nen
    ue = rng.normal(0, 4e-3, m)
    fint, tan, out, states, y = oracle.solid_response(x, E, NU, SY, H, ue)
    assert y and (states[:, 6] > 0).any()
    h = 1e-9
    for i in range(m):
        dp, dm = ue.copy(), ue.copy()
        dp[i] += h
        dm[i] -= h
        num = (oracle.solid_response(x, E, NU, SY, H, dp)[0]
               - oracle.solid_response(x, E, NU, SY, H, dm)[0]) / (2 * h)
        assert np.abs(tan[:, i] - num).max() < 1e-5 * np.abs(tan).max()
    small = ue * 1e-4
    fint, tan, out, states, y = oracle.solid_response(x, E, NU, SY, H, small)
    ke = oracle.solid_stiffness(x, E, NU)
    assert not y
    assert np.abs(tan - ke).max() < 1e-9 * np.abs(ke).max()
    assert np.abs(fint - ke @ small).max() < 1e-9 * np.abs(ke @ small).max()
    ref = oracle.solid_stresses(x, E, NU, small)
    assert np.abs(out[:, :3] - ref[:, :3]).max() == 0.0
    assert np.abs(out[:, 3:] - ref[:, 3:]).max() < 1e-9 * np.abs(ref[:, 3:]).max()


# ------------------------------------------------------------------ the Newton path
def solid_meshes():
    """tests/plasticity.rs:48-64: box_grid([3, 2, 2], [L, W, D]) in Tet4 and Tet10."""
    nodes, tets = M.box_grid((3, 2, 2), (L, W, D))
    qn, qt = M.to_quadratic_tet(nodes, tets)
    return [("Tet4", nodes, tets), ("Tet10", qn, qt)]


def near(a, b):
    return abs(a - b) < 1e-9 * L


def tension_bcs(nodes, delta):
    """tests/plasticity.rs:128-145: symmetry on x = 0, y = 0, z = 0, u_x = delta on x = L."""
    fixed = np.zeros(3 * len(nodes), dtype=np.uint8)
    imposed = np.zeros(3 * len(nodes))
    for i, n in enumerate(nodes):
        for c in range(3):
            if near(n[c], 0.0):
                fixed[3 * i + c] = 1
        if near(n[0], L):
            fixed[3 * i] = 1
            imposed[3 * i] = delta
    return fixed, imposed


def uniaxial_closed_form(strain, hardening):
    """tests/plasticity.rs:147-155."""
    if strain <= SY / E:
        sigma = E * strain
    else:
        sigma = SY + E * hardening / (E + hardening) * (strain - SY / E)
    return sigma, strain - sigma / E


@pytest.mark.parametrize("hardening", [H, 0.0])
@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_uniaxial_tension_loads_and_unloads_like_the_closed_form(mesh, hardening):
    """tests/plasticity.rs:157-228 (8 increments: load to t = 1, unload to zero force at t = 2)."""
    name, nodes, conn = solid_meshes()[mesh]
    strain_max = 3.0 * SY / E
    sigma_max, p_max = uniaxial_closed_form(strain_max, hardening)
    last = p_max / strain_max
    fixed, imposed = tension_bcs(nodes, strain_max * L)
    run = oracle.quasi_static(nodes, conn, E, NU, SY, hardening, fixed, imposed,
                              np.zeros(3 * len(nodes)), 2.0, 8,
                              [(0.0, 0.0), (1.0, 1.0), (2.0, last)], iterative=False)
    assert run["status"] == oracle.OK
    corner = [i for i, n in enumerate(nodes) if near(n[0], L) and near(n[1], W) and near(n[2], D)][0]
    end = [i for i, n in enumerate(nodes) if near(n[0], L)]
    for k in range(9):
        t = 0.25 * k
        lam = t if t <= 1.0 else 1.0 + (last - 1.0) * (t - 1.0)
        strain = lam * strain_max
        sigma, p = (uniaxial_closed_form(strain, hardening) if t <= 1.0
                    else (sigma_max - E * (strain_max - strain), p_max))
        force = sum(run["reactions"][k][3 * i] for i in end)
        assert abs(force - sigma * W * D) < 1e-6 * SY, (name, t, force)
        s = run["stresses"][k]
        assert np.abs(s[..., 3] - sigma).max() < 1e-6 * SY
        assert np.abs(s[..., 4:]).max() < 1e-6 * SY
        assert np.abs(run["p"][k] - p).max() < 1e-9
        uy = run["u"][k][3 * corner + 1]
        assert abs(uy - (-(NU * sigma / E + p / 2.0) * W)) < 1e-6 * strain_max * L
    assert abs(run["reactions"][8][0::3].sum()) < 1e-6 * SY


def uniform(strain, n):
    """tests/plasticity.rs:230-238."""
    return [strain[0] * n[0] + strain[5] / 2 * n[1] + strain[4] / 2 * n[2],
            strain[5] / 2 * n[0] + strain[1] * n[1] + strain[3] / 2 * n[2],
            strain[4] / 2 * n[0] + strain[3] / 2 * n[1] + strain[2] * n[2]]


def boundary_history(nodes, strain):
    fixed = np.zeros(3 * len(nodes), dtype=np.uint8)
    imposed = np.zeros(3 * len(nodes))
    for i, n in enumerate(nodes):
        if near(n[0], 0) or near(n[0], L) or near(n[1], 0) or near(n[1], W) or near(n[2], 0) \
                or near(n[2], D):
            fixed[3 * i:3 * i + 3] = 1
            imposed[3 * i:3 * i + 3] = uniform(strain, n)
    return fixed, imposed


SOLID_STRAIN = [4e-3, -1e-3, -1.5e-3, 1e-3, -0.5e-3, 2e-3]


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_homogeneous_strain_history_matches_the_material_point(mesh):
    """tests/plasticity.rs:239-321 (load, then unload beyond zero: reverse yielding)."""
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = boundary_history(nodes, SOLID_STRAIN)
    run = oracle.quasi_static(nodes, conn, E, NU, SY, H, fixed, imposed,
                              np.zeros(3 * len(nodes)), 2.0, 8,
                              [(0.0, 0.0), (1.0, 1.0), (2.0, -0.5)], iterative=False)
    assert run["status"] == oracle.OK
    state = np.zeros(7)
    for k in range(1, 9):
        t = 0.25 * k
        lam = t if t <= 1.0 else 1.0 - 1.5 * (t - 1.0)
        s, state, _, _ = oracle.von_mises(E, NU, SY, H, np.array(SOLID_STRAIN) * lam, state)
        assert np.abs(run["stresses"][k][..., 3:] - s).max() < 1e-6 * SY, (name, t)
        assert np.abs(run["p"][k] - state[6]).max() < 1e-9
    assert state[6] > 0


@pytest.mark.parametrize("mesh", [0, 1], ids=["Tet4", "Tet10"])
def test_an_elastic_model_gives_the_linear_solution_in_one_iteration(mesh):
    """tests/plasticity.rs:322-370: force-driven, elastic material, two increments of a ramp."""
    name, nodes, conn = solid_meshes()[mesh]
    fixed, imposed = tension_bcs(nodes, 0.0)
    for i, n in enumerate(nodes):
        if near(n[0], L):
            fixed[3 * i] = 0
    f = np.zeros(3 * len(nodes))
    for i, n in enumerate(nodes):
        if near(n[0], L) and near(n[1], W):
            f[3 * i] = 10.0
    prob = M.Problem(name, nodes, conn, np.full(len(conn), E), np.full(len(conn), NU),
                     fixed.astype(bool), imposed, f)
    linear = oracle.model_solve(prob, iterative=False)
    run = oracle.quasi_static(nodes, conn, E, NU, np.inf, 0.0, fixed, imposed, f, 1.0, 2,
                              [(0.0, 0.0), (1.0, 1.0)], iterative=False)
    assert run["status"] == oracle.OK
    for k in (1, 2):
        assert len(run["residuals"][k - 1]) == 2
        assert np.abs(run["u"][k] - 0.5 * k * linear["displacements"]).max() < 1e-9 * 1e-3
    assert np.abs(run["reactions"][2] - linear["reactions"]).max() < 1e-6
    assert np.abs(P.von_mises_of(run["stresses"][2]) - P.von_mises_of(linear["stresses"])).max() < 1e-6
    assert np.abs(run["p"][2]).max() < 1e-15


def test_iterative_newton_matches_the_direct_one_and_converges_quadratically():
    """tests/iterative.rs:165-197 on a Tet10 strip: the PCG and skyline Newton runs take the same
    Newton iterations and agree to 1e-7; tests/plasticity.rs:371-419: once the residual is below
    1e-2, each iteration squares it (bending of a clamped Tet4 bar by an imposed tip
    displacement, a growing plastic zone)."""
    nodes, conn = M.box_grid((6, 2, 2), (2 * L, W, D))
    qn, qt = M.to_quadratic_tet(nodes, conn)
    for nn, cc in ((nodes, conn), (qn, qt)):
        fixed = np.zeros(3 * len(nn), dtype=np.uint8)
        imposed = np.zeros(3 * len(nn))
        for i, n in enumerate(nn):
            if near(n[0], 0.0):
                fixed[3 * i:3 * i + 3] = 1
            elif near(n[0], 2 * L):
                fixed[3 * i + 2] = 1
                imposed[3 * i + 2] = 0.05
        runs = [oracle.quasi_static(nn, cc, E, NU, SY, H, fixed, imposed, np.zeros(3 * len(nn)),
                                    1.0, 5, [(0.0, 0.0), (1.0, 1.0)], iterative=it)
                for it in (False, True)]
        assert all(r["status"] == oracle.OK for r in runs)
        longest = 0
        for k in range(5):
            d, i = runs[0]["residuals"][k], runs[1]["residuals"][k]
            assert len(d) == len(i)
            longest = max(longest, len(d))
            for a, b in zip(d[1:], d[2:]):
                if a < 1e-2:
                    assert b <= max(10 * a * a, 1e-12), d
        gap = np.linalg.norm(runs[1]["u"][-1] - runs[0]["u"][-1]) / np.linalg.norm(runs[0]["u"][-1])
        assert gap < 1e-7
        assert longest >= 4
        assert runs[1]["p"][-1].max() > 0


def test_a_force_beyond_the_limit_load_is_a_typed_error():
    """tests/plasticity.rs:420-462 restated on tetrahedra: a perfectly plastic bar pulled by a
    force above its limit load ends on PlasticCollapse (or, when the cut increments still
    cannot converge, NotConverged), never on a solution."""
    nodes, conn = M.box_grid((4, 1, 1), (L, W, D))
    fixed, imposed = tension_bcs(nodes, 0.0)
    end = [i for i, n in enumerate(nodes) if near(n[0], L)]
    for i in end:
        fixed[3 * i] = 0
    f = np.zeros(3 * len(nodes))
    total = 1.2 * SY * W * D
    for i in end:
        f[3 * i] = total / len(end)
    run = oracle.quasi_static(nodes, conn, E, NU, SY, 0.0, fixed, imposed, f, 1.0, 6,
                              [(0.0, 0.0), (1.0, 1.0)], iterative=False)
    assert run["status"] in (oracle.PLASTIC_COLLAPSE, oracle.NEWTON_NOT_CONVERGED)
    assert run["fail_increment"] >= 5 and run["cuts"] > 0
