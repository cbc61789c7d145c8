"""The benchmark configurations at their stated sizes on the device, against the oracle:

  2. Tet10 cantilever 300x60x60 mm, 60x12x12 cells (225,000 free DOF): FP64 and the mixed
     modes against the oracle's PCG and against the oracle's FP64 skyline Cholesky (the
     reference's direct solver; its solution is the committed fixture
     tests/golden/cantilever_tet10_60x12x12_skyline.npz), with the reference's own metric
     relative_gap < 1e-8 of PCG against the direct solve (tests/iterative.rs:21-32,69-83);
  4a. 256 seeded traction load cases of the Tet10 beam 40x8x8 (69,360 free DOF) solved
     together, against one-at-a-time device solves, the oracle's PCG and its skyline on a sample;
  4b. 64 Tet4 boxes of 500-20,000 DOF in one batch, against the oracle's PCG for every box;
  5. the SIMP box 80x40x40 (403,440 free DOF): the displacements against the oracle's PCG, the
     adjoint gradient against the oracle's FP64 CPU adjoint, and central differences on 32
     random elements within 1e-6, each computed without cancellation.
"""
import os
import time

import numpy as np
import pytest

import oracle
from maille_amd import meshes as M
from conftest import rel_l2, relative_gap

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(f"[configs] {msg}", flush=True)


# ------------------------------------------------------------------ config 2
@pytest.fixture(scope="module")
def tet10():
    t = time.time()
    p = M.cantilever_tet10()
    ref = oracle.model_solve(p, iterative=True)
    assert ref["status"] == oracle.OK
    sky = np.load(os.path.join(GOLDEN, "cantilever_tet10_60x12x12_skyline.npz"))
    assert int(sky["nfree"]) == p.nfree == 225000
    say(f"Tet10 60x12x12: oracle PCG {ref['iterations']} iterations in {time.time() - t:.1f} s")
    return p, ref, sky["displacements"]


def test_tet10_cantilever_oracle_pcg_against_the_skyline(tet10):
    """The oracle's own PCG against its skyline Cholesky at this size: the reference's
    acceptance of its iterative solver (relative_gap < 1e-8, tests/iterative.rs:69-83)."""
    p, ref, u_sky = tet10
    gap = relative_gap(ref["displacements"], u_sky)
    say(f"oracle PCG vs skyline: relative gap {gap:.2e}")
    assert gap < 1e-8


@pytest.mark.parametrize("precision,bound", [("FP64", 1e-10), ("MIXED_IC0", 1e-8),
                                             ("MIXED_JACOBI", 1e-8), ("MIXED", 1e-8)])
def test_tet10_cantilever_on_the_device(gpu, tet10, precision, bound):
    p, ref, u_sky = tet10
    t = time.time()
    with gpu.Model(p, precision=getattr(gpu, precision)) as m:
        if precision == "MIXED_JACOBI":
            m.autotune()
        u, r, info = m.solve()
        sweeps = m.sweeps()
    err = rel_l2(u, ref["displacements"])
    gap = relative_gap(u, u_sky)
    say(f"{precision} (sweeps {sweeps}): {info['iterations']} iterations (oracle "
        f"{ref['iterations']}), rel L2 vs oracle PCG {err:.2e}, relative gap vs skyline "
        f"{gap:.2e}, retries {info['device_retries']} ({time.time() - t:.1f} s)")
    assert err <= bound
    assert gap < 1e-8
    assert info["device_retries"] == 0
    # the mixed factor is applied with FP64 vectors and sums: it costs no more iterations
    if precision in ("FP64", "MIXED_IC0"):
        assert abs(info["iterations"] - ref["iterations"]) <= max(3, 0.12 * ref["iterations"])
    f = ref["f"]
    balance = np.abs(r.reshape(-1, 3).sum(axis=0) + f.reshape(-1, 3).sum(axis=0)).max()
    assert balance <= 1e-9 * np.abs(f.reshape(-1, 3).sum(axis=0)).max()


# ------------------------------------------------------------------ config 4a
BEAM_SAMPLE = (0, 85, 170, 255)


@pytest.fixture(scope="module")
def beam():
    p, F = M.beam_load_cases()
    assert F.shape[0] == 256 and p.nfree == 69360
    s = oracle.System(p.nodes, p.conn, p.young, p.nu, p.fixed)
    assert s.factorize() == oracle.OK
    d = oracle.System(p.nodes, p.conn, p.young, p.nu, p.fixed, iterative=False)
    assert d.factorize() == oracle.OK
    refs = {}
    for k in BEAM_SAMPLE:
        st, u_ref, it, _ = s.solve(F[k])
        assert st == oracle.OK
        st, u_sky, _, _ = d.solve(F[k])
        assert st == oracle.OK
        refs[k] = (u_ref, it, u_sky)
    return p, F, refs


@pytest.mark.parametrize("precision,bound", [("FP64", 1e-10), ("MIXED_JACOBI", 1e-8)])
def test_256_load_cases_of_the_tet10_beam(gpu, beam, precision, bound):
    """The 256 cases as one multi-right-hand-side block (one PCG per case in shared launches):
    FP64 with replicated exact triangular solves, MIXED_JACOBI with the FP32 factor's Jacobi
    sweeps over every case at once; against the oracle's PCG and skyline on four cases and the
    same model's one-at-a-time solves."""
    p, F, refs = beam
    with gpu.Model(p, precision=getattr(gpu, precision)) as m:
        m.factor()
        m.autotune(nrhs=len(F))
        t = time.time()
        U, infos = m.solve_many(F)
        say(f"beam {precision} (sweeps {m.sweeps()}): 256 cases solved together in "
            f"{time.time() - t:.1f} s, iterations {min(i['iterations'] for i in infos)}-"
            f"{max(i['iterations'] for i in infos)}")
        seq = {k: m.solve(F[k], reactions=False)[0] for k in BEAM_SAMPLE}
    for k in BEAM_SAMPLE:
        u_ref, it, u_sky = refs[k]
        err, gap = rel_l2(U[k], u_ref), relative_gap(U[k], u_sky)
        say(f"case {k}: {infos[k]['iterations']} iterations (oracle {it}), rel L2 vs oracle "
            f"PCG {err:.2e}, vs the sequential device solve {rel_l2(U[k], seq[k]):.2e}, "
            f"relative gap vs skyline {gap:.2e}")
        assert err <= bound
        assert rel_l2(U[k], seq[k]) <= (1e-10 if precision == "FP64" else 1e-9)
        assert gap < 1e-8
    assert all(i["device_retries"] == 0 for i in infos)


# ------------------------------------------------------------------ config 4b
@pytest.fixture(scope="module")
def boxes():
    bs = M.box_batch()
    t = time.time()
    refs = [oracle.model_solve(b, iterative=True) for b in bs]
    assert all(r["status"] == oracle.OK for r in refs)
    say(f"64 boxes: oracle PCG solves in {time.time() - t:.1f} s")
    return bs, refs


@pytest.mark.parametrize("precision,bound", [("FP64", 1e-10), ("MIXED_IC0", 1e-8),
                                             ("MIXED_JACOBI", 1e-8)])
def test_64_uneven_boxes_in_one_batch(gpu, boxes, precision, bound):
    boxes, refs = boxes
    sizes = [b.ndof for b in boxes]
    assert len(boxes) == 64 and min(sizes) >= 400 and max(sizes) <= 24000
    models = [gpu.Model(b) for b in boxes]
    try:
        with gpu.Batch(models, precision=getattr(gpu, precision)) as batch:
            t = time.time()
            us, infos = batch.solve()
            say(f"{precision}: 64 boxes ({sum(b.nfree for b in boxes)} free DOF) in "
                f"{time.time() - t:.2f} s")
        worst = 0.0
        for b, ref, u, info in zip(boxes, refs, us, infos):
            worst = max(worst, rel_l2(u, ref["displacements"]))
            assert info["device_retries"] == 0
        say(f"{precision}: worst rel L2 vs the oracle's PCG {worst:.2e}")
        assert worst <= bound
    finally:
        for m in models:
            m.close()


# ------------------------------------------------------------------ config 5
@pytest.fixture(scope="module")
def simp(gpu):
    p = M.simp_box()
    assert p.nfree == 403440
    m = gpu.Model(p)
    u, _, info = m.solve(reactions=False)
    g = m.compliance_gradient()
    yield p, m, u, g, info
    m.close()


def test_simp_box_solution_and_adjoint_against_the_oracle(simp):
    p, m, u, g, info = simp
    t = time.time()
    ref = oracle.model_solve(p, iterative=True)
    assert ref["status"] == oracle.OK
    err = rel_l2(u, ref["displacements"])
    say(f"SIMP: {info['iterations']} iterations (oracle {ref['iterations']}), rel L2 {err:.2e} "
        f"({time.time() - t:.1f} s)")
    assert err <= 1e-10
    # the device adjoint and the oracle's FP64 CPU adjoint: bit for bit at the same u, and
    # within the solves' agreement at the oracle's own u
    assert np.array_equal(g, oracle.compliance_gradient(p.nodes, p.conn, p.young, p.nu, u))
    g_ref = oracle.compliance_gradient(p.nodes, p.conn, p.young, p.nu, ref["displacements"])
    assert rel_l2(g, g_ref) <= 1e-9
    # compliance is homogeneous of degree -1 in the moduli: C = -sum_e E_e dC/dE_e
    c = float(p.f @ u)
    assert abs(c + float(p.young @ g)) <= 1e-9 * abs(c)


def test_simp_box_central_differences_on_32_elements(gpu, simp):
    """Central differences of C = f.u on 32 random elements (seed 6), computed without
    cancellation: with K+- = K +- h k_e (k_e: the element stiffness at unit modulus),
    C(E_e + h) - C(E_e - h) = f.(u+ - u-) = -2h u+.k_e u- exactly (K+ symmetric), so the
    central quotient is -u+.k_e u-, a product of element-local quantities of the two perturbed
    solutions. It must agree with the adjoint within 1e-6 (relative step 1e-4: the quotient's
    truncation error is second order)."""
    p, m, u, g, info = simp
    rng = np.random.default_rng(6)
    elems = rng.choice(len(p.conn), 32, replace=False)
    probe = gpu.Model(p, precision=gpu.MIXED_JACOBI)
    try:
        probe.autotune()
        t = time.time()
        worst = 0.0
        for e in elems:
            h = 1e-4 * p.young[e]
            sols = []
            for sgn in (1.0, -1.0):
                young = p.young.copy()
                young[e] += sgn * h
                probe.set_young(young)
                probe.assemble()
                probe.factor()
                ue, _, inf = probe.solve(reactions=False)
                assert inf["device_retries"] == 0
                sols.append(ue)
            dofs = (3 * p.conn[e][:, None] + np.arange(3)).ravel()
            ke = oracle.solid_stiffness(p.nodes[p.conn[e]], 1.0, p.nu[e])
            fd = -float(sols[0][dofs] @ ke @ sols[1][dofs])
            rel = abs(fd - g[e]) / abs(g[e])
            worst = max(worst, rel)
            assert rel <= 1e-6, f"element {e}: adjoint {g[e]:.12e} central {fd:.12e}"
        say(f"SIMP central differences on 32 elements: worst relative difference {worst:.2e} "
            f"({time.time() - t:.1f} s, sweeps {probe.sweeps()})")
    finally:
        probe.close()
