"""Device parity of the transient heat-conduction path (Model::solve_step, model.rs:620-706)
against the oracle's restatement, on the reference's transient scenarios (tests/transient.rs).
The element matrices, the left matrix's assembly, the carried term, the factor and the
triangular solves are the oracle's operations in the oracle's order; the PCG dot products are
tree sums, so temperatures agree to round-off: within 1e-10 of the oracle's PCG steps (relative
to the temperature range), reactions and heat fluxes within 1e-8 (relative). The thermal loads
(heat sources, convection) are host code and equal the oracle's bit for bit."""
import numpy as np
import pytest

import oracle
from maille_amd import thermal as T
from test_transient_oracle import (K, L, RHO_C, convection_at, fix_plane, fourier, mean, slab,
                                   walled, weights)

pytestmark = pytest.mark.gpu


def compare(run, incs, scale):
    assert run["status"] == oracle.OK
    for k, inc in enumerate(incs):
        assert np.abs(inc.temperatures - run["temperatures"][k]).max() <= 1e-10 * scale, k
        rs = max(np.abs(run["reactions"][k]).max(), 1e-300)
        assert np.abs(inc.reactions - run["reactions"][k]).max() <= 1e-8 * max(rs, scale), k
        assert np.array_equal(inc.heat_fluxes[..., :3], run["fluxes"][k][..., :3])
        qs = max(np.abs(run["fluxes"][k][..., 3:]).max(), 1e-300)
        assert np.abs(inc.heat_fluxes[..., 3:] - run["fluxes"][k][..., 3:]).max() <= 1e-8 * qs, k


@pytest.mark.parametrize("name", ["Tet4", "Tet10"])
def test_walled_slab_with_sources_matches_the_oracle(gpu, name):
    """tests/transient.rs:178-233: from zero towards the steady state (implicit Euler), and from
    the steady state with Crank-Nicolson, which stays there."""
    nodes, conn, fixed, imposed, f = walled(name)
    sources = [(e, 300.0) for e in range(len(conn))]
    m = T.ThermalModel(nodes, conn, K, RHO_C, fixed, imposed, heat_sources=sources)
    assert np.array_equal(m.assemble_loads(), f)
    incs, info = m.solve_step(T.Step(4.0, 8), np.zeros(len(nodes)))
    run = oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, f, np.zeros(len(nodes)), 4.0, 8)
    compare(run, incs, 100.0)
    assert info["iterations"] > 0
    st, steady, sreact = oracle.thermal_steady(nodes, conn, K, fixed, imposed, f)
    incs, _ = m.solve_step(T.Step(1.0, 5).with_theta(0.5), steady)
    for inc in incs[1:]:
        assert np.abs(inc.temperatures - steady).max() < 1e-10
        assert np.abs(inc.reactions - sreact).max() < 1e-9


@pytest.mark.parametrize("name,nx", [("Tet4", 40), ("Tet10", 10)])
def test_quenched_slab_follows_the_fourier_series_on_the_device(gpu, name, nx):
    """tests/transient.rs:250-283 on the device (400 implicit Euler steps)."""
    ts = 100.0
    nodes, conn = slab(name, nx)
    fixed, imposed = fix_plane(nodes, 0.0, ts)
    fix_plane(nodes, L, ts, fixed, imposed)
    zero = np.zeros(len(nodes))
    m = T.ThermalModel(nodes, conn, K, RHO_C, fixed, imposed)
    incs, _ = m.solve_step(T.Step(0.1, 400), zero)
    run = oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, zero, zero, 0.1, 400)
    compare(run, incs, ts)
    for s in (80, 200, 400):
        err = np.abs(incs[s].temperatures - fourier(nodes[:, 0], incs[s].time, ts)).max() / ts
        assert err < 4e-3


@pytest.mark.parametrize("name", ["Tet4", "Tet10"])
def test_small_biot_cooling_on_the_device(gpu, name):
    """tests/transient.rs:301-337 on the device: convection faces in K and F, no fixed node,
    three values of theta reusing one model (refactored per step size and theta)."""
    film, sink, t0 = 0.4, 20.0, 100.0
    tau = RHO_C * (L * 0.25 * 0.25) / (film * 0.25 * 0.25)
    nodes, conn = slab(name, 2)
    w = weights(nodes, conn)
    conv = convection_at(nodes, conn, L, film, sink)
    zero = np.zeros(len(nodes))
    fixed = np.zeros(len(nodes), dtype=np.uint8)
    m = T.ThermalModel(nodes, conn, 1e7 * film * L, RHO_C, fixed, zero, convection=conv)
    f = oracle.thermal_loads(nodes, conn, zero, (), conv)
    assert np.array_equal(m.assemble_loads(), f)
    for theta in (1.0, 0.75, 0.5):
        incs, _ = m.solve_step(T.Step(tau, 8).with_theta(theta), np.full(len(nodes), t0))
        run = oracle.transient(nodes, conn, 1e7 * film * L, RHO_C, fixed, zero, f,
                               np.full(len(nodes), t0), tau, 8, theta=theta, convection=conv)
        # k is 1e7 times the film coefficient: the mean temperature lies in the near-null space
        # of K, where a 1e-12 relative residual leaves ~1e-9 relative freedom; the device and
        # oracle PCGs (dot products summed in different orders) both stop inside it
        for k, inc in enumerate(incs):
            assert np.abs(inc.temperatures - run["temperatures"][k]).max() <= 1e-7 * (t0 - sink)
        g = (1.0 - (1.0 - theta) / 8) / (1.0 + theta / 8)
        for s, inc in enumerate(incs):
            assert abs(mean(inc.temperatures, w) - (sink + (t0 - sink) * g ** s)) < 1e-5 * (t0 - sink)


def test_convection_and_sources_on_a_tet10_slab_match_the_oracle(gpu):
    """A Tet10 slab with a fixed face, convection on the opposite face and a local source, theta
    = 0.75: the device steps equal the oracle's PCG steps."""
    nodes, conn = slab("Tet10", 6)
    fixed, imposed = fix_plane(nodes, 0.0, 100.0)
    conv = convection_at(nodes, conn, L, 2.0, 10.0)
    zero = np.zeros(len(nodes))
    m = T.ThermalModel(nodes, conn, K, RHO_C, fixed, imposed, heat_sources=[(0, 50.0)],
                       convection=conv)
    f = oracle.thermal_loads(nodes, conn, zero, [(0, 50.0)], conv)
    assert np.array_equal(m.assemble_loads(), f)
    incs, _ = m.solve_step(T.Step(0.5, 10).with_theta(0.75), zero)
    run = oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, f, zero, 0.5, 10, theta=0.75,
                           convection=conv)
    compare(run, incs, 100.0)


@pytest.mark.parametrize("name", ["Tet4", "Tet10"])
def test_steady_thermal_solve_matches_the_oracle(gpu, name):
    """Model::solve of the thermal model on the device: the walled slab with sources and a
    convection face, against the oracle's steady solve; a following Step from that field stays
    on it (the reuse of one handle for both)."""
    nodes, conn, fixed, imposed, f = walled(name)
    conv = convection_at(nodes, conn, L, 3.0, 15.0)
    fixed2 = fixed.copy()
    fixed2[np.abs(nodes[:, 0] - L) < 1e-12] = 0
    sources = [(e, 300.0) for e in range(len(conn))]
    m = T.ThermalModel(nodes, conn, K, RHO_C, fixed2, imposed, heat_sources=sources,
                       convection=conv)
    f2 = oracle.thermal_loads(nodes, conn, np.zeros(len(nodes)), sources, conv)
    inc, info = m.solve()
    st, t_ref, r_ref = oracle.thermal_steady(nodes, conn, K, fixed2, imposed, f2, convection=conv)
    assert st == oracle.OK
    assert np.abs(inc.temperatures - t_ref).max() <= 1e-10 * np.abs(t_ref).max()
    assert np.abs(inc.reactions - r_ref).max() <= 1e-8 * np.abs(r_ref).max()
    assert info["iterations"] > 0
    incs, _ = m.solve_step(T.Step(1.0, 4).with_theta(0.5), inc.temperatures)
    for i in incs[1:]:
        assert np.abs(i.temperatures - inc.temperatures).max() < 1e-9 * np.abs(t_ref).max()


@pytest.mark.parametrize("name", ["Tet4", "Tet10"])
def test_jacobi_sweeps_apply_to_the_transient_caller(gpu, name):
    """The thermal handle with the FP32 factor applied by Jacobi sweeps (ML_MIXED_JACOBI): its
    one-DOF-per-node pattern, whose rows hold at most 16 strict entries (Tet4), starts on the
    narrowest fused variant; every step still meets PCG's 1e-12 residual, so the temperatures
    stay within 1e-8 of the oracle's FP64 steps (relative to the range) and the reactions within
    1e-6."""
    import maille_amd as ma
    nodes, conn, fixed, imposed, f = walled(name)
    sources = [(e, 300.0) for e in range(len(conn))]
    m = T.ThermalModel(nodes, conn, K, RHO_C, fixed, imposed, heat_sources=sources,
                       precision=ma.MIXED_JACOBI)
    incs, info = m.solve_step(T.Step(4.0, 8), np.zeros(len(nodes)), fluxes=False)
    run = oracle.transient(nodes, conn, K, RHO_C, fixed, imposed, f, np.zeros(len(nodes)), 4.0, 8)
    assert run["status"] == oracle.OK
    for k, inc in enumerate(incs):
        assert np.abs(inc.temperatures - run["temperatures"][k]).max() <= 1e-8 * 100.0, k
        rs = max(np.abs(run["reactions"][k]).max(), 1e-300)
        assert np.abs(inc.reactions - run["reactions"][k]).max() <= 1e-6 * max(rs, 100.0), k
    assert info["iterations"] > 0
