"""Parity of the HIP path with the oracle (the FP64 restatement of the reference), through the
C ABI. Bit-exact where the device restates the reference's arithmetic operation for operation
(element stiffness, assembled K_ff, SpMV, IC(0) factor, triangular solves, reactions,
stresses); displacements of the PCG solve within 1e-10 relative L2 of the oracle's PCG (FP64)
and 1e-8 (mixed), the parity tolerances of the solve path."""
import dataclasses

import numpy as np
import pytest

import oracle
from maille_amd import meshes as M
from conftest import rel_l2, relative_gap

pytestmark = pytest.mark.gpu


def _problems():
    return {
        "reference_cantilever": M.reference_cantilever(),
        "cantilever_tet4": M.cantilever_tet4(),
        "jittered_cube_8": M.jittered_cube(8),
        "cantilever_tet10_small": M.cantilever_tet10((12, 3, 3), (60.0, 15.0, 15.0)),
    }


PROBLEMS = _problems()


def _oracle_system(p):
    s = oracle.System(p.nodes, p.conn, p.young, p.nu, p.fixed, p.imposed)
    return s


@pytest.mark.parametrize("nen", [4, 10])
def test_stiffness_batch_is_bit_exact(gpu, nen):
    rng = np.random.default_rng(7)
    nodes, tets = M.box_grid((3, 2, 2), (3.0, 2.0, 2.0))
    if nen == 10:
        nodes, tets = M.to_quadratic_tet(nodes, tets)
    coords = nodes[tets]
    young = rng.uniform(5e4, 2.5e5, len(tets))
    nu = rng.uniform(0.1, 0.45, len(tets))
    print(f"stiffness batch nen={nen} ne={len(tets)}", flush=True)
    ke = gpu.stiffness_batch(coords, young, nu)
    for e in range(len(tets)):
        ref = oracle.solid_stiffness(coords[e], young[e], nu[e])
        assert np.array_equal(ke[e], ref), f"element {e}"


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_assembled_matrix_is_bit_exact(gpu, name):
    p = PROBLEMS[name]
    print(f"assemble {name} ndof={p.ndof}", flush=True)
    m = gpu.Model(p)
    s = _oracle_system(p)
    order, off, cols, vals = m.csr()
    ooff, ocols, ovals = s.csr()
    assert np.array_equal(order, s.order())
    assert np.array_equal(off, ooff) and np.array_equal(cols, ocols)
    assert np.array_equal(vals, ovals)


@pytest.mark.parametrize("name", ["reference_cantilever", "jittered_cube_8"])
def test_spmv_factor_and_preconditioner(gpu, name):
    p = PROBLEMS[name]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    print(f"pcg operator {name} n={s.n}", flush=True)
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    csr = oracle.Csr.from_arrays(off, cols, vals)
    x = np.sin(0.37 * np.arange(s.n))
    assert np.array_equal(pcg.multiply(x), csr.multiply(x))  # bit-exact SpMV
    assert s.factorize() == oracle.OK
    l, shift = pcg.factor()
    assert shift == s.shift_index
    assert np.array_equal(l, s.l_values())  # bit-exact IC(0)
    r = np.cos(0.11 * np.arange(s.n))
    assert np.array_equal(pcg.precondition(r), csr.precondition(s.l_values(), r))


@pytest.mark.parametrize("precision", ["MIXED", "MIXED_IC0", "MIXED_JACOBI"])
def test_mixed_preconditioner_applies_the_fp32_factor(gpu, precision):
    """A mixed handle stores its factor in FP32 and applies it with FP64 vectors and sums: its
    exact preconditioner is, bit for bit, the oracle's FP64 precondition (iterative.rs:
    173-188) of the FP32 factor widened, on every call."""
    p = PROBLEMS["jittered_cube_8"]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    print(f"{precision} preconditioner n={s.n}", flush=True)
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals), precision=getattr(gpu, precision))
    pcg.set_sweeps(0)
    r = np.cos(0.11 * np.arange(s.n))
    z = pcg.precondition(r)
    l, _ = pcg.factor()  # the FP32 factor, widened
    assert np.array_equal(l.astype(np.float32).astype(np.float64), l)
    csr = oracle.Csr.from_arrays(off, cols, vals)
    assert np.array_equal(z, csr.precondition(l, r))
    assert np.array_equal(pcg.precondition(r), z)
    # the FP32 factor is the FP64 one to FP32 accuracy (it is factored in FP32 arithmetic)
    assert s.factorize() == oracle.OK
    l64 = s.l_values()
    assert np.abs(l - l64).max() <= 1e-5 * np.abs(l64).max()


SWEEP_CASES = [("jittered_cube_8", k, f, 1.0) for k in (1, 2, 3, 4, 7) for f in (1, 116, 216, 224, 0)] + \
    [("cantilever_tet10_small", k, f, 1.0) for k in (1, 3) for f in (1, 116, 148, 0)] + \
    [("jittered_cube_8", 3, f, 0.8) for f in (1, 116, 216, 0)] + \
    [("cantilever_tet10_small", k, f, w) for k in (2, 4) for f in (1, 148, 0) for w in (0.8, 0.6)]


@pytest.mark.parametrize("name,sweeps,fused,weight", SWEEP_CASES)
def test_jacobi_sweeps_match_their_restatement(gpu, name, sweeps, fused, weight):
    """ML_MIXED_JACOBI's preconditioner: `sweeps` Jacobi sweeps of L y = r and of L^T z = y on
    the FP32 factor, bit for bit the float64 restatement, on every call, whether all sweeps of
    a system run in one launch (fused 1: one row per lane with every row in registers, 24
    coefficients on the cube and 80 on the Tet10 mesh, whose rows reach 146 and re-read their
    tails; 116 and 148: variants that keep fewer and re-read more; 216, 224: two rows per lane
    in blocks of 512 rows) or one launch per sweep (0), undamped or damped by a weight; more
    sweeps approach the exact triangular solves."""
    from sweeps import jacobi_precondition
    p = PROBLEMS[name]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    print(f"jacobi {name} sweeps={sweeps} fused={fused} weight={weight} n={s.n}", flush=True)
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals), precision=gpu.MIXED_JACOBI)
    assert pcg.sweeps() == 4 and pcg.sweep_weight() == 1.0
    pcg.set_sweeps(sweeps)
    pcg.set_sweep_weight(weight)
    pcg.set_sweeps_fused(fused)
    assert pcg.sweeps() == sweeps
    code = {1: 124 if name == "jittered_cube_8" else 180}.get(fused, fused)
    assert pcg.sweeps_fused() == code
    r = np.cos(0.11 * np.arange(s.n))
    z = pcg.precondition(r)
    l, _ = pcg.factor()
    assert np.array_equal(z, jacobi_precondition(off, cols, l, r, sweeps, weight))
    assert np.array_equal(pcg.precondition(r), z)
    exact = oracle.Csr.from_arrays(off, cols, vals).precondition(l, r)
    far = rel_l2(jacobi_precondition(off, cols, l, r, 1), exact)
    assert rel_l2(jacobi_precondition(off, cols, l, r, 12), exact) < far


def test_fused_and_per_launch_sweeps_solve_identically(gpu):
    """The two forms of the sweeps give the same preconditioner bit for bit and the same
    per-block partial sums of r.z, so whole MIXED_JACOBI solves agree exactly: the same
    iterates, iteration counts and residuals."""
    p = PROBLEMS["jittered_cube_8"]
    print("fused vs per-launch solve", flush=True)
    ref = oracle.model_solve(p, iterative=True)
    for weight in (1.0, 0.8):
        out = {}
        with gpu.Model(p, precision=gpu.MIXED_JACOBI) as m:
            m.set_sweeps(3)
            m.set_sweep_weight(weight)
            for fused in (1, 116, 224, 0):
                m.set_sweeps_fused(fused)
                assert m.sweeps_fused() == {1: 124}.get(fused, fused)
                u, r, info = m.solve()
                assert info["device_retries"] == 0
                out[fused] = (u, r, info)
        u1, r1, i1 = out[0]
        for fused in (1, 116, 224):
            u2, r2, i2 = out[fused]
            assert np.array_equal(u1, u2) and np.array_equal(r1, r2)
            assert i1["iterations"] == i2["iterations"] and i1["residual"] == i2["residual"]
        assert rel_l2(u1, ref["displacements"]) <= 1e-8


def test_fp64_handle_refuses_sweeps(gpu):
    p = PROBLEMS["reference_cantilever"]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    pcg.set_sweeps(0)
    with pytest.raises(ValueError, match="exactly"):
        pcg.set_sweeps(2)
    with pytest.raises(ValueError, match="no sweeps"):
        pcg.set_sweep_weight(0.8)
    mixed = gpu.Pcg(gpu.Csr(off, cols, vals), precision=gpu.MIXED_JACOBI)
    for bad in (0.0, -0.5, 1.5):
        with pytest.raises(ValueError, match="weight"):
            mixed.set_sweep_weight(bad)


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_fp64_solve_matches_oracle_pcg(gpu, name):
    p = PROBLEMS[name]
    print(f"fp64 solve {name}", flush=True)
    ref = oracle.model_solve(p, iterative=True)
    assert ref["status"] == oracle.OK
    m = gpu.Model(p)
    u, r, info = m.solve()
    assert rel_l2(u, ref["displacements"]) <= 1e-10
    # only the dot products are summed in another order (per wave, then across waves): the
    # iteration count stays close; how close depends on how long the residual lingers near
    # the tolerance (up to ~10% on the cantilever with an imposed displacement)
    assert abs(info["iterations"] - ref["iterations"]) <= max(3, 0.12 * ref["iterations"])
    # the recursive residual of PCG drifts from b - Ax; the device drifts like the oracle
    assert info["true_residual"] <= 10 * ref["true_residual"] + 1e-12
    assert relative_gap(r, ref["reactions"]) < 1e-8
    st = m.stresses(u)
    assert relative_gap(st[..., 3:], ref["stresses"][..., 3:]) < 1e-7
    assert np.array_equal(st[..., :3], ref["stresses"][..., :3])


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_fp64_solve_with_reference_dots_is_the_oracle_bit_for_bit(gpu, name):
    """With the PCG's dot products in the reference's eight-lane order (k_dot8, solve.rs:
    100-110), every operation of the FP64 Pcg::solve follows the reference's: the device takes
    the oracle's iteration count and returns its displacements and reactions bit for bit."""
    p = PROBLEMS[name]
    print(f"reference-dot solve {name}", flush=True)
    ref = oracle.model_solve(p, iterative=True)
    assert ref["status"] == oracle.OK
    with gpu.Model(p) as m:
        m.set_reference_dots(True)
        u, r, info = m.solve()
        assert info["iterations"] == ref["iterations"]
        assert np.array_equal(u, ref["displacements"])
        assert info["residual"] == ref["residual"]
        # and back to the tree sums: rounding-level agreement only
        m.set_reference_dots(False)
        u2, _, info2 = m.solve()
        assert rel_l2(u2, ref["displacements"]) <= 1e-10


def test_pcg_with_reference_dots_is_the_oracle_pcg_bit_for_bit(gpu):
    """Pcg::new + Pcg::solve on an assembled CSR, FP64 with the reference's dot order: the
    oracle's PCG iterate for iterate (solution, iteration count, final residual)."""
    p = PROBLEMS["jittered_cube_8"]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    b = np.sin(0.3 * np.arange(s.n)) + 0.5
    st, x_ref, it_ref, res_ref, shift = oracle.Csr.from_arrays(off, cols, vals).pcg(b)
    assert st == oracle.OK
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    pcg.set_reference_dots(True)
    x, info = pcg.solve(b, info=True)
    assert info["iterations"] == it_ref and info["shift_index"] == shift
    assert np.array_equal(x, x_ref)
    assert info["residual"] == res_ref
    with pytest.raises(ValueError):
        gpu.Pcg(gpu.Csr(off, cols, vals), precision=gpu.MIXED).set_reference_dots(True)


@pytest.mark.parametrize("precision", ["MIXED", "MIXED_IC0", "MIXED_JACOBI"])
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_mixed_solve_matches_oracle(gpu, name, precision):
    """FP32 operator and/or preconditioner with FP64 iterate, residual and dots: within
    1e-8 relative L2 of the oracle's FP64 PCG (the bound for mixed precision)."""
    p = PROBLEMS[name]
    print(f"{precision} solve {name}", flush=True)
    ref = oracle.model_solve(p, iterative=True)
    m = gpu.Model(p, precision=getattr(gpu, precision))
    u, r, info = m.solve()
    print(f"  iterations {info['iterations']} (oracle {ref['iterations']}), replacements "
          f"{info['refinements']}, escalated {info['escalated']}, true residual "
          f"{info['true_residual']:.1e}", flush=True)
    assert rel_l2(u, ref["displacements"]) <= 1e-8


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_stresses_and_nodal_stresses_are_bit_exact(gpu, name):
    """The Gauss-point stresses (element.rs:270-302) and the node-averaged stresses
    (results.rs:138-166,184-221) of one displacement field, on the device and in the oracle,
    bit for bit."""
    p = PROBLEMS[name]
    print(f"stresses {name}", flush=True)
    m = gpu.Model(p)
    u = np.random.default_rng(11).standard_normal(m.ndof) * 1e-3
    gp_ref = oracle.stresses(p.nodes, p.conn, p.young, p.nu, u)
    assert np.array_equal(m.stresses(u), gp_ref)
    assert np.array_equal(m.nodal_stresses(u), oracle.nodal_stresses(p.nodes, p.conn, gp_ref))


def test_iterative_matches_direct_on_reference_cantilever(gpu):
    """tests/iterative.rs:69-83 with the device PCG against the oracle's skyline Cholesky."""
    p = PROBLEMS["reference_cantilever"]
    direct = oracle.model_solve(p, iterative=False)
    u, r, _ = gpu.Model(p).solve()
    assert relative_gap(u, direct["displacements"]) < 1e-8
    assert relative_gap(r, direct["reactions"]) < 1e-8


def test_shift_ladder_on_the_device(gpu):
    """iterative.rs:201-205: Kershaw's matrix fails unshifted IC(0); the device climbs the
    ladder to the same shift as the oracle and solves."""
    from test_oracle import KERSHAW, kershaw_csr
    a = kershaw_csr(oracle)
    st, x_ref, _, _, shift = a.pcg(np.array([1.0, 2.0, 3.0, 4.0]))
    off, cols, vals = a.arrays()
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    l, dshift = pcg.factor()
    assert dshift == shift > 0
    x = pcg.solve(np.array([1.0, 2.0, 3.0, 4.0]))
    assert np.abs(KERSHAW @ x - [1, 2, 3, 4]).max() < 1e-9


def test_singular_and_zero_right_hand_side(gpu):
    """iterative.rs:211-216 (zero b gives exactly x = 0) and tests/iterative.rs:200-224 (an
    unconstrained structure is reported as Singular or IterativeNotConverged)."""
    p = PROBLEMS["cantilever_tet4"]
    m = gpu.Model(p)
    u, _, info = m.solve(np.zeros(p.ndof))
    assert np.all(u == 0) and info["iterations"] == 0
    q = M.reference_cantilever((4, 1, 1))
    q.fixed[:] = False
    q.f[:] = 0.0
    q.f[2] = 1.0
    with pytest.raises((gpu.Singular, gpu.IterativeNotConverged)):
        gpu.Model(q).solve()


def test_pressure_loads_match_oracle(gpu):
    p = M.jittered_cube(6)
    m = gpu.Model(p)
    f = m.loads()
    g = oracle.add_pressures(p.nodes, p.conn, p.pressures, p.f.copy())
    assert np.array_equal(f, g)


@pytest.mark.parametrize("name", ["jittered_cube_8", "cantilever_tet10_small"])
def test_edge_loads_match_oracle(gpu, name):
    """edge_load (assembly.rs:680-710) on the edges of the x = max end, added after the
    pressures as Model::assemble does (model.rs:1113-1127): the library's F equals the oracle's
    bit for bit, and the FP64 solve under it stays within 1e-10 of the oracle's PCG."""
    base = PROBLEMS[name]
    xmax = base.nodes[:, 0].max()
    loads = []
    for e, row in enumerate(base.conn):
        for k, (a, b) in enumerate([(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]):
            if abs(base.nodes[row[a], 0] - xmax) < 1e-9 and abs(base.nodes[row[b], 0] - xmax) < 1e-9:
                loads.append((e, k, (0.5, -1.25, 2.0)))
    assert loads
    p = dataclasses.replace(base, edge_loads=loads)
    print(f"edge loads {name}: {len(loads)}", flush=True)
    m = gpu.Model(p)
    f = m.loads()
    g = oracle.add_edge_loads(p.nodes, p.conn, loads,
                              oracle.add_pressures(p.nodes, p.conn, p.pressures, p.f.copy()))
    assert np.array_equal(f, g)
    u, _, _ = m.solve()
    ref = oracle.model_solve(p, iterative=True)
    assert rel_l2(u, ref["displacements"]) <= 1e-10


BEAM_TOML = """
units = "N, mm, MPa"
mesh = "beam.msh"
[materials.alu]
young = 70e3
poisson = 0.33
[[sections]]
set = "solide"
material = "alu"
[[bcs]]
set = "encastrement"
dof = "ux"
value = 0.0
[[bcs]]
set = "encastrement"
dof = "uy"
value = 0.0
[[bcs]]
set = "encastrement"
dof = "uz"
value = 0.0
[[pressures]]
set = "bout"
value = 2.5
[[edge_loads]]
set = "arete"
value = [0.0, -4.0, 1.0]
[[loads]]
set = "coin"
dof = "uz"
value = -50.0
"""


def test_model_files_solve_on_the_device(gpu, tmp_path):
    """A model file and its .msh read and built by the library (maille_amd.model_file,
    build.rs:27-235) and solved on the device: the reference's example cube in traction gives
    the analytic field (tests/model.rs:194-228), and a Tet10 beam written Gmsh's way (half the
    cells reversed, shuffled tags, three volume blocks) with a clamped face, an end pressure,
    a line load on an end edge and a point load matches the oracle's solve of the same Problem
    within 1e-10."""
    import os
    import shutil

    import msh_cases
    from maille_amd import model_file as mf
    from test_model_file import CUBE

    golden = os.path.join(os.path.dirname(__file__), "golden", "cube.msh")
    shutil.copy(golden, tmp_path / "cube.msh")
    (tmp_path / "cube.toml").write_text(CUBE)
    print("model file: cube.toml", flush=True)
    p, u, r, info = mf.solve_file(str(tmp_path / "cube.toml"))
    u = u.reshape(-1, 3)
    x = p.nodes
    assert np.abs(u[:, 0] - 0.01 * x[:, 0]).max() < 1e-12
    assert np.abs(u[:, 1] + 0.25 * 0.01 * x[:, 1]).max() < 1e-12
    assert np.abs(u[:, 2] + 0.25 * 0.01 * x[:, 2]).max() < 1e-12

    beam = M.cantilever_tet10((12, 3, 3), (60.0, 15.0, 15.0))
    nodes, cells = beam.nodes, beam.conn
    face = lambda pred: [tuple(int(v) for v in cells[e][list(M.TET_FACES[f])])
                         for e, f in zip(*M.faces_on(nodes, cells, pred))]
    edge = set()
    for c in cells:
        for a, b in M.TET10_EDGES:
            pa, pb = nodes[c[a]], nodes[c[b]]
            if max(pa[0], pb[0]) == 60.0 and min(pa[0], pb[0]) == 60.0 and pa[1] == pb[1] == 0.0:
                edge.add(tuple(sorted((int(c[a]), int(c[b])))))
    corner = [int(np.argmax((nodes == [60.0, 15.0, 15.0]).all(axis=1)))]
    groups = [("solide", 3, None), ("encastrement", 2, face(lambda q: q[:, 0] == 0.0)),
              ("bout", 2, face(lambda q: q[:, 0] == 60.0)), ("arete", 1, sorted(edge)),
              ("coin", 0, corner)]
    (tmp_path / "beam.msh").write_text(msh_cases.write_msh(nodes, cells, groups, seed=2))
    (tmp_path / "beam.toml").write_text(BEAM_TOML)
    print("model file: beam.toml", flush=True)
    p = mf.load(str(tmp_path / "beam.toml"))
    assert np.array_equal(p.conn, cells) and len(p.pressures) and len(p.edge_loads) == 3
    u, _, info = gpu.Model(p).solve()
    ref = oracle.model_solve(p, iterative=True)
    assert rel_l2(u, ref["displacements"]) <= 1e-10


def test_stalled_dataflow_waits_end_cleanly(gpu):
    """A dataflow launch whose waits expire (forced from the host with a one-tick bound, so
    every wave still ends normally) is repeated on smaller grids and then reported as a
    device timeout; the handle solves correctly again once the bound is restored."""
    p = PROBLEMS["jittered_cube_8"]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    b = np.sin(0.37 * np.arange(s.n))
    x_ref, info_ref = pcg.solve(b, info=True)
    assert info_ref["device_retries"] == 0
    pcg.set_wait_limit(1)
    print("solve with a one-tick wait bound", flush=True)
    with pytest.raises(gpu.DeviceError, match="device wait expired"):
        pcg.solve(b)
    pcg.set_wait_limit(0)
    x, info = pcg.solve(b, info=True)
    assert info["device_retries"] == 0
    assert rel_l2(x, x_ref) <= 1e-10


def test_compliance_gradient_matches_cpu_adjoint(gpu):
    """The device adjoint dC/dE_e (one thread per element over the stored K_e) is bit-identical
    to the oracle's FP64 CPU adjoint at the same displacements, and within 1e-9 of it at the
    oracle's own solution."""
    p = M.simp_box((16, 8, 8), (16.0, 8.0, 8.0))
    m = gpu.Model(p)
    u, _, _ = m.solve()
    g = m.compliance_gradient()
    assert np.array_equal(g, oracle.compliance_gradient(p.nodes, p.conn, p.young, p.nu, u))
    ref = oracle.model_solve(p, iterative=True)
    g_ref = oracle.compliance_gradient(p.nodes, p.conn, p.young, p.nu, ref["displacements"])
    assert rel_l2(g, g_ref) <= 1e-9


def test_compliance_gradient_matches_central_differences(gpu):
    """Central differences of C = f.u from device solves after ml_model_set_young agree with the
    device adjoint on random elements within 1e-6. With K+- = K +- h k_e (k_e: the element
    stiffness at unit modulus), C(E_e + h) - C(E_e - h) = f.(u+ - u-) = -2h u+.k_e u- exactly,
    so the central quotient is taken as -u+.k_e u- (relative step 1e-4, truncation error of
    order 1e-8), free of the cancellation of two nearly equal compliances. The plain difference
    of compliances is checked too, at relative step 1e-3, to the 1e-5 its cancellation allows."""
    p = M.simp_box((8, 4, 4), (8.0, 4.0, 4.0))
    m = gpu.Model(p)
    u, _, _ = m.solve()
    g = m.compliance_gradient()
    rng = np.random.default_rng(6)

    def perturbed(e, h):
        sols = []
        for sgn in (1.0, -1.0):
            young = p.young.copy()
            young[e] += sgn * h
            m.set_young(young)
            m.assemble()
            m.factor()
            sols.append(m.solve(reactions=False)[0])
        return sols

    for e in rng.choice(len(p.conn), 6, replace=False):
        dofs = (3 * p.conn[e][:, None] + np.arange(3)).ravel()
        ke = oracle.solid_stiffness(p.nodes[p.conn[e]], 1.0, p.nu[e])
        up, um = perturbed(e, 1e-4 * p.young[e])
        fd = -float(up[dofs] @ ke @ um[dofs])
        h = 1e-3 * p.young[e]
        up, um = perturbed(e, h)
        plain = (float(p.f @ up) - float(p.f @ um)) / (2 * h)
        print(f"element {e}: adjoint {g[e]:.12e} central {fd:.12e} plain {plain:.9e}", flush=True)
        assert abs(fd - g[e]) <= 1e-6 * abs(g[e])
        assert abs(plain - g[e]) <= 1e-5 * abs(g[e])
    imposed = p.imposed.copy()
    imposed[np.flatnonzero(p.fixed)[0]] = 0.01  # a nonzero imposed displacement
    q = M.Problem(p.name, p.nodes, p.conn, p.young, p.nu, p.fixed, imposed, p.f)
    mq = gpu.Model(q)
    mq.solve()
    with pytest.raises(ValueError, match="homogeneous"):
        mq.compliance_gradient()


# ------------------------------------------------------------------ batched solves
def test_many_right_hand_sides_match_separate_solves(gpu):
    """Pcg::solve on several right-hand sides at once (one PCG per right-hand side, every SpMV
    and triangular solve covering all of them) against one solve at a time and the oracle's
    PCG; a zero right-hand side gives exactly x = 0."""
    p = PROBLEMS["jittered_cube_8"]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    csr = oracle.Csr.from_arrays(off, cols, vals)
    rng = np.random.default_rng(11)
    B = rng.normal(size=(5, s.n))
    B[3] = 0.0
    X, info = pcg.solve(B, info=True)
    assert np.all(X[3] == 0.0)
    for k in (0, 1, 2, 4):
        x1 = pcg.solve(B[k])
        assert rel_l2(X[k], x1) <= 1e-10
        st, xo, it, res, _ = csr.pcg(B[k])
        assert st == oracle.OK and rel_l2(X[k], xo) <= 1e-10


def test_load_cases_of_one_model(gpu):
    """Model load cases solved together (the beam of the batched workload, at a smaller size)
    match the model's one-at-a-time solves."""
    p, F = M.beam_load_cases(n=(12, 3, 3), size=(60.0, 15.0, 15.0), cases=6)
    m = gpu.Model(p)
    U, infos = m.solve_many(F)
    assert len(infos) == 6
    for k in range(6):
        u1, _, info1 = m.solve(F[k], reactions=False)
        assert rel_l2(U[k], u1) <= 1e-10
        assert abs(infos[k]["iterations"] - info1["iterations"]) <= 3


@pytest.mark.parametrize("precision,sweeps,lanes,weight", [
    ("FP64", 0, 0, 1.0), ("MIXED_IC0", 0, 0, 1.0), ("MIXED_JACOBI", 1, 0, 1.0),
    ("MIXED_JACOBI", 3, 0, 1.0), ("MIXED_JACOBI", 4, 0, 1.0), ("MIXED_JACOBI", 3, 8, 1.0),
    ("MIXED_JACOBI", 2, 32, 1.0), ("MIXED_JACOBI", 3, 0, 0.8), ("MIXED_JACOBI", 2, 16, 0.6)])
def test_batched_preconditioner_is_the_single_one_bit_for_bit(gpu, precision, sweeps, lanes,
                                                              weight):
    """The preconditioner as batched solves apply it to many right-hand sides (replicated exact
    solves, or the Jacobi sweeps of every vector in one launch per sweep on interleaved copies,
    with the lanes per row following the vectors of a pass or set) gives each vector exactly
    what the single application gives it; the sweeps equal the float64 restatement bit for
    bit. 11 and 40 vectors (the last lane group partial), with a zero vector among them."""
    from sweeps import jacobi_precondition
    p = PROBLEMS["jittered_cube_8"]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    print(f"precondition_many {precision} sweeps={sweeps} n={s.n}", flush=True)
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals), precision=getattr(gpu, precision))
    if precision == "MIXED_JACOBI":
        pcg.set_sweeps(sweeps)
        pcg.set_sweep_weight(weight)
    pcg.set_batch_lanes(lanes)
    assert pcg.batch_lanes() == lanes
    rng = np.random.default_rng(19)
    l, _ = pcg.factor()
    for count in (11, 40):
        R = rng.normal(size=(count, s.n))
        R[6] = 0.0
        Z = pcg.precondition_many(R)
        for k in range(len(R)):
            assert np.array_equal(Z[k], pcg.precondition(R[k]))
            if precision == "MIXED_JACOBI" and k < 12:
                assert np.array_equal(Z[k], jacobi_precondition(off, cols, l, R[k], sweeps,
                                                                weight))
        assert np.all(Z[6] == 0.0)


@pytest.mark.parametrize("name", ["jittered_cube_8", "cantilever_tet10_small"])
def test_batched_spmv_is_the_single_one_bit_for_bit(gpu, name):
    """The SpMV of many vectors at once (interleaved copies, a group of lanes per row and up to
    64 vectors) equals the single SpMV of each vector, and the oracle's Csr::multiply, bit for
    bit; 1, 5, 11, 40 and 70 vectors cover every lane-group width."""
    p = PROBLEMS[name]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    print(f"multiply_many {name} n={s.n}", flush=True)
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    csr = oracle.Csr.from_arrays(off, cols, vals)
    rng = np.random.default_rng(23)
    for k in (1, 5, 11, 40, 70):
        X = rng.normal(size=(k, s.n))
        Y = pcg.multiply_many(X)
        for v in range(k):
            y1 = pcg.multiply(X[v])
            assert np.array_equal(Y[v], y1)
            if v < 2:
                assert np.array_equal(Y[v], csr.multiply(X[v]))


def test_batched_width_from_the_autotuner_keeps_results(gpu):
    """The autotuner's lanes per row of the batched SpMV and sweeps (and its right-hand sides
    per pass) change no result: each case's arithmetic is the same at every width, so the
    batched solve after autotuning equals the one before bit for bit (sweep count pinned)."""
    p, F = M.beam_load_cases(n=(12, 3, 3), size=(60.0, 15.0, 15.0), cases=20)
    print("batched width from the autotuner", flush=True)
    with gpu.Model(p, precision=gpu.MIXED_JACOBI) as m:
        m.set_sweeps(2)
        assert m.batch_lanes() == 0
        U1, i1 = m.solve_many(F)
        m.autotune(nrhs=len(F))
        assert m.batch_lanes() in (8, 16, 32) and m.sweeps() == 2
        U2, i2 = m.solve_many(F)
    assert np.array_equal(U1, U2)
    assert [a["iterations"] for a in i1] == [b["iterations"] for b in i2]


def test_load_cases_with_jacobi_sweeps(gpu):
    """Load cases of one MIXED_JACOBI model solved together (the preconditioner's sweeps over
    every case at once) against the model's one-at-a-time solves and the oracle's FP64 PCG."""
    p, F = M.beam_load_cases(n=(12, 3, 3), size=(60.0, 15.0, 15.0), cases=10)
    print("load cases with Jacobi sweeps", flush=True)
    with gpu.Model(p, precision=gpu.MIXED_JACOBI) as m:
        m.set_sweeps(2)
        U, infos = m.solve_many(F)
        assert len(infos) == 10 and all(i["device_retries"] == 0 for i in infos)
        for k in (0, 3, 9):
            u1, _, info1 = m.solve(F[k], reactions=False)
            assert rel_l2(U[k], u1) <= 1e-9
            assert abs(infos[k]["iterations"] - info1["iterations"]) <= 3
            ref = oracle.model_solve(dataclasses.replace(p, f=F[k].copy()), iterative=True)
            assert rel_l2(U[k], ref["displacements"]) <= 1e-8


@pytest.mark.parametrize("precision", ["FP64", "MIXED_IC0", "MIXED_JACOBI"])
def test_batch_of_uneven_models(gpu, precision):
    """Boxes of uneven sizes and materials solved in one batch (block-diagonal operator, one
    PCG per model) against each model solved alone and against the oracle's PCG."""
    boxes = M.box_batch(count=6, lo=300.0, hi=3000.0)
    models = [gpu.Model(b) for b in boxes]
    batch = gpu.Batch(models, precision=getattr(gpu, precision))
    us, infos = batch.solve()
    bound = 1e-10 if precision == "FP64" else 1e-8
    for b, m, u, info in zip(boxes, models, us, infos):
        ref = oracle.model_solve(b, iterative=True)
        print(f"{b.name}: {b.nfree} free DOF, {info['iterations']} iterations "
              f"(oracle {ref['iterations']}), shift {info['shift_index']}", flush=True)
        assert rel_l2(u, ref["displacements"]) <= bound
        assert info["shift_index"] == ref["shift_index"] or precision != "FP64"
        u1, _, _ = m.solve(reactions=False)
        assert rel_l2(u, u1) <= bound


def test_autotuner_keeps_results(gpu):
    """The autotuner only changes schedules: the preconditioner and the solution stay bit for
    bit the same after it has picked its grids."""
    p = PROBLEMS["jittered_cube_8"]
    s = _oracle_system(p)
    off, cols, vals = s.csr()
    pcg = gpu.Pcg(gpu.Csr(off, cols, vals))
    r = np.cos(0.11 * np.arange(s.n))
    b = np.sin(0.37 * np.arange(s.n))
    z0, x0 = pcg.precondition(r), pcg.solve(b)
    waves, per_pass, occ = pcg.autotune(nrhs=16)
    print(f"autotuned: {waves} waves per CU (occupancy {occ}), {per_pass} right-hand sides per pass",
          flush=True)
    assert 1 <= waves <= occ and per_pass in (8, 16, 32, 64, 128, 256)
    assert np.array_equal(pcg.precondition(r), z0)
    assert np.array_equal(pcg.solve(b), x0)


def test_a_model_with_every_dof_imposed(gpu):
    """Model::solve when every DOF is imposed (FreeSystem::new with no unknowns): u is the
    imposed field and the reactions K u - F equal the oracle's bit for bit."""
    p = M.cantilever_tet4((4, 2, 2), (20.0, 10.0, 10.0))
    strain = np.array([1e-3, -2e-4, 3e-4, 1e-4, -5e-4, 2e-4])
    x = p.nodes
    u0 = np.stack([strain[0] * x[:, 0] + strain[5] / 2 * x[:, 1] + strain[4] / 2 * x[:, 2],
                   strain[5] / 2 * x[:, 0] + strain[1] * x[:, 1] + strain[3] / 2 * x[:, 2],
                   strain[4] / 2 * x[:, 0] + strain[3] / 2 * x[:, 1] + strain[2] * x[:, 2]],
                  axis=1).ravel()
    q = dataclasses.replace(p, fixed=np.ones(p.ndof, dtype=bool), imposed=u0, pressures=[])
    m = gpu.Model(q)
    u, r, info = m.solve()
    assert np.array_equal(u, u0) and info["iterations"] == 0
    ref = oracle.reactions(q.nodes, q.conn, q.young, q.nu, q.fixed, q.f, u0)
    assert np.array_equal(r, ref)


def test_dataflow_grids_are_bounded_by_a_co_residency_probe(gpu):
    """Each dataflow kernel's grid comes from probes of the kernel itself (every workgroup of
    three probes in a row resident at once), one workgroup per CU below that, hence below the
    runtime's occupancy figure; solves on those grids need no retry."""
    p = PROBLEMS["jittered_cube_8"]
    for precision in ("FP64", "MIXED_IC0", "MIXED_JACOBI"):
        with gpu.Model(p, precision=getattr(gpu, precision)) as m:
            res = m.residency()
            print(f"{precision}: reported/probed workgroups per CU {res}", flush=True)
            for kernel, (reported, used) in res.items():
                if kernel == "fused_sweeps" and precision != "MIXED_JACOBI":
                    assert (reported, used) == (0, 0)  # no sweep layouts: nothing probed
                    continue
                assert 1 <= used and (used < reported or used == 1), kernel
            u, _, info = m.solve()
            assert info["device_retries"] == 0
