"""The benchmark cube at its full size on the device (jittered Tet4 cube 100^3: 6,000,000
elements, 3,060,300 free DOF): the assembled K_ff, the SpMV, the IC(0) factor and the
preconditioner application equal the oracle's bit for bit, and the solves meet size-independent
properties (true residual, balance of reactions against the applied load, mixed precision
against FP64). The displacements are compared with the oracle's OpenMP PCG (to rounding, its sums
run in another order), and, with the PCG's dot products in the reference's order, with the
oracle's serial PCG bit for bit."""
import time

import numpy as np
import pytest

import oracle
from maille_amd import meshes as M
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[full size] {msg}", flush=True)


@pytest.fixture(scope="module")
def cube():
    t = time.time()
    p = M.jittered_cube(100)
    say(f"mesh {len(p.conn)} elements in {time.time() - t:.1f} s")
    return p


@pytest.fixture(scope="module")
def cube_oracle(cube):
    p = cube
    t = time.time()
    s = oracle.System(p.nodes, p.conn, p.young, p.nu, p.fixed, p.imposed)
    assert s.factorize() == oracle.OK
    say(f"oracle system n={s.n} in {time.time() - t:.1f} s")
    return s


def test_full_size_operator_factor_and_preconditioner_are_bit_exact(gpu, cube, cube_oracle):
    p = cube
    s = cube_oracle
    off, cols, vals = s.csr()
    t = time.time()
    t = time.time()
    m = gpu.Model(p)
    order, doff, dcols, dvals = m.csr()
    say(f"device model in {time.time() - t:.1f} s")
    assert np.array_equal(order, s.order())
    assert np.array_equal(doff, off) and np.array_equal(dcols, cols)
    assert np.array_equal(dvals, vals)  # bit-exact assembly of 69.3M entries
    del m, order, doff, dcols, dvals
    t = time.time()
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    a = oracle.Csr.from_arrays(off, cols, vals)
    x = np.sin(0.37 * np.arange(s.n))
    assert np.array_equal(pcg.multiply(x), a.multiply(x))  # bit-exact SpMV
    l, shift = pcg.factor()
    assert shift == s.shift_index
    lv = s.l_values()
    assert np.array_equal(l, lv)  # bit-exact IC(0)
    r = np.cos(0.11 * np.arange(s.n))
    assert np.array_equal(pcg.precondition(r), a.precondition(lv, r))
    say(f"SpMV, factor and preconditioner compared in {time.time() - t:.1f} s")


def test_full_size_solutions_meet_the_size_independent_properties(gpu, cube):
    p = cube
    t = time.time()
    m = gpu.Model(p)
    f = m.loads()
    u, r, info = m.solve(f)
    say(f"FP64 solve: {info['iterations']} iterations, true residual {info['true_residual']:.2e} "
        f"in {time.time() - t:.1f} s")
    assert info["true_residual"] <= 1e-10
    # equilibrium: the reactions on the clamped face carry the applied pressure load
    total_f = f.reshape(-1, 3).sum(axis=0)
    total_r = r.reshape(-1, 3).sum(axis=0)
    assert np.abs(total_f + total_r).max() <= 1e-9 * np.abs(total_f).max()
    assert np.abs(r[~p.fixed]).max() == 0.0
    # the load is a unit pressure on the unit top face: 1 N downwards in all
    assert abs(total_f[2] + 1.0) < 1e-12 and np.abs(total_f[:2]).max() < 1e-12
    t = time.time()
    mm = gpu.Model(p, precision=gpu.MIXED_IC0)
    um, _, im = mm.solve(f, reactions=False)
    say(f"mixed_ic0 solve: {im['iterations']} iterations in {time.time() - t:.1f} s")
    assert rel_l2(um, u) <= 1e-8
    # compression: every top node moves down
    top = np.abs(p.nodes[:, 2] - 1.0) < 1e-12
    assert (u.reshape(-1, 3)[top, 2] < 0).all()


def test_full_size_displacements_match_the_oracle_pcg(gpu, cube, cube_oracle):
    """The benchmark cube's displacements against the oracle's PCG at full size (its OpenMP
    form, whose sums run in another order than the serial restatement, so agreement is to
    rounding): FP64 within 1e-10 (relative L2), the mixed modes within 1e-8."""
    p = cube
    s = cube_oracle
    f = p.f.copy()
    oracle.add_pressures(p.nodes, p.conn, p.pressures, f)
    t = time.time()
    st, u_ref, it_ref, res = s.solve_parallel(f, threads=16)
    assert st == oracle.OK
    say(f"oracle OpenMP PCG: {it_ref} iterations, |r|/|b| {res:.1e} in {time.time() - t:.1f} s")
    for name, bound in (("FP64", 1e-10), ("MIXED_IC0", 1e-8), ("MIXED_JACOBI", 1e-8)):
        t = time.time()
        with gpu.Model(p, precision=getattr(gpu, name)) as m:
            u, _, info = m.solve(f, reactions=False)
        err = rel_l2(u, u_ref)
        say(f"{name}: {info['iterations']} iterations, rel L2 vs oracle {err:.2e}, "
            f"retries {info['device_retries']} ({time.time() - t:.1f} s)")
        assert err <= bound
        assert info["device_retries"] == 0
        if name == "FP64":
            assert abs(info["iterations"] - it_ref) <= max(3, 0.12 * it_ref)


def test_full_size_fp64_solve_with_reference_dots_is_the_oracle_bit_for_bit(gpu, cube,
                                                                            cube_oracle):
    """The whole FP64 Pcg::solve of the benchmark cube with the PCG's dot products in the
    reference's eight-lane order: the device takes the oracle's serial PCG's iteration count and
    returns its displacements bit for bit at 3,060,300 free DOF (the oracle's serial solve takes
    about a minute on one host core)."""
    p = cube
    s = cube_oracle
    f = p.f.copy()
    oracle.add_pressures(p.nodes, p.conn, p.pressures, f)
    t = time.time()
    st, u_ref, it_ref, res_ref = s.solve(f)
    assert st == oracle.OK
    say(f"oracle serial PCG: {it_ref} iterations in {time.time() - t:.1f} s")
    t = time.time()
    with gpu.Model(p) as m:
        m.set_reference_dots(True)
        u, _, info = m.solve(f, reactions=False)
    say(f"device FP64 with reference dots: {info['iterations']} iterations in "
        f"{time.time() - t:.1f} s, identical: {np.array_equal(u, u_ref)}")
    assert info["iterations"] == it_ref
    assert info["residual"] == res_ref
    assert np.array_equal(u, u_ref)
