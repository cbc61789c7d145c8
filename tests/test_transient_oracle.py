"""Pins the oracle's restatement of the transient heat-conduction path (Model::solve_step, the
theta method on tetrahedra) without running the reference: the reference's own transient tests
restated for Tet4 and Tet10 (tests/transient.rs), closed forms of the element matrices, and the
Fourier series and lumped-cooling amplification factor those tests compare with."""
import numpy as np
import pytest

import oracle
from maille_amd import meshes as M

K, RHO, CP, L, H, THICKNESS = 1.0, 2.0, 0.5, 1.0, 0.25, 0.25
RHO_C = RHO * CP
VOLUME = L * H * THICKNESS
TET_FACES = ((0, 1, 2), (0, 3, 1), (1, 3, 2), (2, 3, 0))


def slab(name, nx):
    """tests/transient.rs:28-72: L x H x THICKNESS, nx elements along x, one across."""
    nodes, tets = M.box_grid((nx, 1, 1), (L, H, THICKNESS))
    if name == "Tet10":
        nodes, tets = M.to_quadratic_tet(nodes, tets)
    return nodes, tets


def fix_plane(nodes, x, value, fixed=None, imposed=None):
    """tests/transient.rs:79-91."""
    fixed = np.zeros(len(nodes), dtype=np.uint8) if fixed is None else fixed
    imposed = np.zeros(len(nodes)) if imposed is None else imposed
    on = np.abs(nodes[:, 0] - x) < 1e-12
    fixed[on] = 1
    imposed[on] = value
    return fixed, imposed


def convection_at(nodes, conn, x, film, sink):
    """tests/transient.rs:93-110: every element face lying in the plane x."""
    out = []
    for e, t in enumerate(conn):
        for face, verts in enumerate(TET_FACES):
            if all(abs(nodes[t[v], 0] - x) < 1e-12 for v in verts):
                out.append((e, face, film, sink))
    return out


def weights(nodes, conn):
    """tests/transient.rs:121-151: rho c int N_i dV per node, without the oracle: V/4 at each
    Tet4 vertex; -V/20 at Tet10 vertices and V/5 at mid-nodes."""
    w = np.zeros(len(nodes))
    vol = M.signed_volume(nodes, conn) / 6.0
    for e, t in enumerate(conn):
        if len(t) == 4:
            w[t] += RHO_C * vol[e] / 4.0
        else:
            w[t[:4]] += RHO_C * vol[e] * -0.05
            w[t[4:]] += RHO_C * vol[e] * 0.2
    return w


def mean(temps, w):
    return float(temps @ w) / (RHO_C * VOLUME)


@pytest.mark.parametrize("nen", [4, 10])
def test_element_matrices_closed_forms(nen):
    """Conduction rows sum to zero and reproduce a linear field's flux; capacity entries sum to
    rho c V and its row sums are rho c int N_i; the source gives q int N_i."""
    rng = np.random.default_rng(3)
    x4 = np.array([[0, 0, 0], [1.2, 0.1, 0], [0.2, 0.9, 0.1], [0.1, 0.2, 1.1]]) + rng.normal(0, .03, (4, 3))
    x = x4 if nen == 4 else np.vstack([x4, [(x4[a] + x4[b]) / 2 for a, b in M.TET10_EDGES]])
    vol = M.signed_volume(x4, np.array([[0, 1, 2, 3]]))[0] / 6.0
    k, rc, q = 2.5, 3.0, 7.0
    ke, ce, fe = oracle.thermal_element(x, k, rc, q)
    assert np.abs(ke.sum(axis=1)).max() < 1e-12 * np.abs(ke).max()
    assert np.abs(ke - ke.T).max() < 1e-14 * np.abs(ke).max()
    grad = np.array([0.3, -1.1, 0.7])
    t = x @ grad
    # K t = int k grad N . grad T = boundary flux terms; for the total: sum_i (K t)_i = 0 and
    # t' K t = k V |grad|^2
    assert abs(t @ ke @ t - k * vol * grad @ grad) < 1e-12 * k * vol * (grad @ grad)
    assert abs(ce.sum() - rc * vol) < 1e-13 * rc * vol
    expect = (np.full(4, vol / 4) if nen == 4
              else np.r_[np.full(4, -vol / 20), np.full(6, vol / 5)])
    assert np.abs(ce.sum(axis=1) - rc * expect).max() < 1e-13 * rc * vol
    assert np.abs(fe - q * expect).max() < 1e-13 * q * vol


def test_increments_start_from_the_initial_field_at_regular_times():
    """tests/transient.rs:154-176 (on Tet4)."""
    nodes, conn = slab("Tet4", 4)
    initial = 10.0 * nodes[:, 0]
    fixed, imposed = fix_plane(nodes, 0.0, 30.0)
    run = oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, np.zeros(len(nodes)), initial,
                           0.5, 4)
    assert run["status"] == oracle.OK and run["temperatures"].shape == (5, len(nodes))
    expected = np.where(np.abs(nodes[:, 0]) < 1e-12, 30.0, 10.0 * nodes[:, 0])
    assert np.abs(run["temperatures"][0] - expected).max() < 1e-15
    assert np.abs(run["reactions"][0]).max() < 1e-15


def walled(name):
    nodes, conn = slab(name, 4)
    fixed, imposed = fix_plane(nodes, 0.0, 100.0)
    fix_plane(nodes, L, 20.0, fixed, imposed)
    f = oracle.thermal_loads(nodes, conn, np.zeros(len(nodes)), [(e, 300.0) for e in range(len(conn))])
    return nodes, conn, fixed, imposed, f


@pytest.mark.parametrize("name", ["Tet4", "Tet10"])
def test_starting_from_the_steady_state_stays_there(name):
    """tests/transient.rs:178-204: the steady state is a fixed point of the theta method."""
    nodes, conn, fixed, imposed, f = walled(name)
    st, steady, sreact = oracle.thermal_steady(nodes, conn, K, fixed, imposed, f)
    assert st == oracle.OK
    run = oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, f, steady, 1.0, 5, theta=0.5)
    for k in range(1, 6):
        assert np.abs(run["temperatures"][k] - steady).max() < 1e-10
        assert np.abs(run["reactions"][k] - sreact).max() < 1e-9


@pytest.mark.parametrize("name", ["Tet4", "Tet10"])
def test_long_time_reaches_the_steady_state(name):
    """tests/transient.rs:206-233."""
    nodes, conn, fixed, imposed, f = walled(name)
    _, steady, _ = oracle.thermal_steady(nodes, conn, K, fixed, imposed, f)
    run = oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, f, np.zeros(len(nodes)), 40.0, 40)
    assert np.abs(run["temperatures"][40] - steady).max() < 1e-8


def fourier(x, t, ts):
    """tests/transient.rs:235-248."""
    alpha = K / RHO_C
    n = 2 * np.arange(500) + 1.0
    s = (4.0 / (n * np.pi) * np.sin(n * np.pi * x[:, None] / L)
         * np.exp(-alpha * (n * np.pi / L) ** 2 * t)).sum(axis=1)
    return ts * (1.0 - s)


def quench_errors(name, nx, increments):
    """tests/transient.rs:250-267."""
    ts = 100.0
    nodes, conn = slab(name, nx)
    fixed, imposed = fix_plane(nodes, 0.0, ts)
    fix_plane(nodes, L, ts, fixed, imposed)
    run = oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, np.zeros(len(nodes)),
                           np.zeros(len(nodes)), 0.1, increments)
    dt = 0.1 / increments
    return [float(np.abs(run["temperatures"][s] - fourier(nodes[:, 0], s * dt, ts)).max() / ts)
            for s in (increments // 5, increments // 2, increments)]


def test_quenched_slab_follows_the_fourier_series():
    """tests/transient.rs:269-283 (Tet4 40 elements, Tet10 10, 400 implicit Euler steps)."""
    for name, nx in (("Tet4", 40), ("Tet10", 10)):
        assert all(e < 4e-3 for e in quench_errors(name, nx, 400)), name


def test_quenched_slab_converges_in_time_and_space():
    """tests/transient.rs:285-299."""
    ratio = quench_errors("Tet10", 20, 100)[2] / quench_errors("Tet10", 20, 400)[2]
    assert 3.4 < ratio < 4.4
    ratio = quench_errors("Tet4", 10, 400)[2] / quench_errors("Tet4", 20, 400)[2]
    assert 2.9 < ratio < 4.2


@pytest.mark.parametrize("name", ["Tet4", "Tet10"])
def test_small_biot_cooling_follows_the_discrete_amplification_factor(name):
    """tests/transient.rs:301-337: near-infinite conductivity, convection on x = L: the mean
    temperature follows g = (1 - (1 - theta) r) / (1 + theta r) per step."""
    film, sink, t0 = 0.4, 20.0, 100.0
    tau = RHO_C * VOLUME / (film * H * THICKNESS)
    nodes, conn = slab(name, 2)
    w = weights(nodes, conn)
    conv = convection_at(nodes, conn, L, film, sink)
    fixed = np.zeros(len(nodes), dtype=np.uint8)
    f = oracle.thermal_loads(nodes, conn, np.zeros(len(nodes)), (), conv)
    for theta in (1.0, 0.75, 0.5):
        run = oracle.transient(nodes, conn, 1e7 * film * L, RHO_C, fixed, np.zeros(len(nodes)), f,
                               np.full(len(nodes), t0), tau, 8, theta=theta, convection=conv)
        assert run["status"] == oracle.OK
        r = (tau / 8) / tau
        g = (1.0 - (1.0 - theta) * r) / (1.0 + theta * r)
        for s in range(9):
            got = mean(run["temperatures"][s], w)
            assert abs(got - (sink + (t0 - sink) * g ** s)) < 1e-5 * (t0 - sink), (theta, s)
            assert np.abs(run["temperatures"][s] - got).max() < 1e-5 * (t0 - sink)


def test_iterative_transient_matches_the_direct_one():
    """tests/iterative.rs:94-121 restated on a Tet10 slab: PCG and skyline steps agree."""
    nodes, conn = slab("Tet10", 6)
    fixed, imposed = fix_plane(nodes, 0.0, 100.0)
    conv = convection_at(nodes, conn, L, 2.0, 10.0)
    f = oracle.thermal_loads(nodes, conn, np.zeros(len(nodes)), [(0, 50.0)], conv)
    runs = [oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, f, np.zeros(len(nodes)), 0.5,
                             10, theta=0.75, convection=conv, iterative=it) for it in (False, True)]
    gap = np.abs(runs[1]["temperatures"] - runs[0]["temperatures"]).max() / 100.0
    assert gap < 1e-9
    assert np.abs(runs[1]["fluxes"] - runs[0]["fluxes"]).max() < 1e-7 * np.abs(runs[0]["fluxes"]).max()
