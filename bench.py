"""Benchmark: the 3.1M-DOF Tet4 linear-elastic solve on one MI355X.

Workload: unit cube of 100^3 Kuhn cubes x 6 Tet4 (6M elements, 1,030,301 nodes, 3,060,300
free DOF), interior nodes jittered by up to 0.2h (seed 1), per-element E lognormal (median
200000, sigma 0.5, seed 2), nu = 0.3, clamped at z = 0, 1 MPa pressure on z = 1.

A step is one device solve with the mesh, pattern and loads resident in HBM: element
stiffness + gather assembly of K_ff, IC(0) factorisation with the shift ladder, PCG to the
reference's 1e-12 relative residual, reactions. Three modes are timed: the FP64 drop-in,
``mixed_ic0`` (FP64 SpMV, FP32 IC(0) applied by exact triangular solves with FP64 vectors) and
``mixed_jacobi`` (the FP32 factor applied by Jacobi sweeps; ``--precision mixed`` adds the FP32
SpMV with residual replacement); the value is the fastest mode whose displacements stay within
1e-8 (relative L2) of the FP64 drop-in.

Prints one JSON line (see the keys below). Kernel durations come from HIP events recorded on
the library's stream around every launch of the kernel class during the timed steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3.1M-DOF Tet4 PCG solve s + SpMV GB/s/HBM peak, mixed vs FP64, rel err 1e-8"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n, nnz_full, val, index_bytes):
    """Algorithmic SpMV bytes: values + the column indices and per-row fields of the form in use
    (int32 or 16-bit offsets per entry with row lengths, or one 16-bit stream per group of rows
    with equal columns with its per-row fields: Model.spmv_index) + x + y."""
    return nnz_full * val + index_bytes + 2 * n * val


def sweep_bytes(n, nnz_strict, col_bytes=4):
    """Algorithmic bytes of one Jacobi sweep (k_sweep): FP32 coefficients and the columns
    (int32, or 16-bit distances from the row) of the strict triangle, row lengths, and the
    FP64 right-hand side, reciprocal diagonal, previous iterate and result."""
    return nnz_strict * (4 + col_bytes) + n * (4 + 8 + 8 + 8 + 8)


def fused_sweep_bytes(n, nnz_strict, sweeps):
    """Algorithmic bytes of one launch of the fused sweeps (k_sweeps_fused: every sweep of one
    triangular system): the FP32 coefficients and 16-bit LDS slots of the strict triangle once,
    the row lengths (1 byte), the FP64 right-hand side, reciprocal diagonal and result once,
    and each of the sweeps - 1 intermediate iterates written once and read once. The staging
    of each block's external columns re-reads values other blocks wrote: it counts as traffic,
    not here."""
    return nnz_strict * (4 + 2) + n * (1 + 8 + 8 + 8) + (sweeps - 1) * n * (8 + 8)


def trisolve_bytes(n, nnz_l, val):
    """Algorithmic bytes of one triangular solve with L (or L^T): entries + row pointers +
    right-hand side + solution."""
    return nnz_l * (val + 4) + (n + 1) * 4 + 2 * n * val


def cpu_baseline(problem, iterations, sample_iters):
    """The oracle (single-threaded FP64 restatement of the reference) on a bounded sample of
    the same workload: full assembly and IC(0) factorisation, then ``sample_iters`` PCG
    iterations, extrapolated to ``iterations`` iterations."""
    import oracle
    f = problem.f.copy()
    oracle.add_pressures(problem.nodes, problem.conn, problem.pressures, f)
    s = oracle.System(problem.nodes, problem.conn, problem.young, problem.nu, problem.fixed)
    st = s.factorize()
    st, u, it, res = s.solve(f, maxit=sample_iters)
    t = s.times()
    per_iter = t["solve"] / max(it, 1)
    est = t["assembly"] + t["factor"] + per_iter * iterations
    # the stronger, labelled baseline: the same PCG iterations with OpenMP on the host cores
    # (SpMV row-parallel, level-scheduled triangular solves); assembly and IC(0) as above
    import oracle as _o
    threads = _o.lib().or_omp_threads()
    par_iter, _ = s.pcg_parallel(f, sample_iters, threads)
    multicore = {
        "value": round(t["assembly"] + t["factor"] + par_iter * iterations, 3), "unit": "s",
        "cores": threads, "kind": "port",
        "sample": (f"the oracle's PCG with OpenMP on {threads} threads, {sample_iters} iterations "
                   f"measured at {par_iter:.4f} s each, extrapolated to {iterations}; assembly "
                   f"{t['assembly']:.2f} s and IC(0) {t['factor']:.2f} s single-threaded"),
    }
    return {
        "value": round(est, 3), "unit": "s", "cores": 1, "kind": "port",
        "sample": (f"oracle on this workload: assembly {t['assembly']:.2f} s + IC(0) "
                   f"{t['factor']:.2f} s measured, {it} PCG iterations measured at "
                   f"{per_iter:.3f} s each, extrapolated to {iterations} iterations"),
        "phases_s": {"assembly": round(t["assembly"], 3), "factor": round(t["factor"], 3),
                     "pcg_iteration": round(per_iter, 4), "pattern": round(t["pattern"], 3)},
        "multicore": multicore,
    }


def emit(out):
    print(json.dumps(out), flush=True)


def run_beam_cases(args, ma, meshes):
    """256 load cases of the Tet10 beam solved together (ml_model_solve_many) against the same
    model's one-at-a-time solves (the reference's sequential Model::solve calls), in FP64
    (replicated exact triangular solves) and mixed_jacobi (the FP32 factor's Jacobi sweeps over
    every case at once); the line's value is the faster mode within 1e-8 of FP64."""
    p, F = meshes.beam_load_cases(cases=args.cases)
    modes, U64 = {}, None
    for name, code in (("fp64", ma.FP64), ("mixed_jacobi", ma.MIXED_JACOBI)):
        with ma.Model(p, precision=code) as m:
            m.factor()
            chosen = m.autotune(nrhs=len(F))
            log(f"beam {name}: {p.nfree} free DOF, {len(F)} cases; autotuned {chosen}, "
                f"{m.sweeps()} sweeps (weight {m.sweep_weight()}), {m.batch_lanes()} lanes per row")
            m.solve_many(F[:8])
            t0 = time.perf_counter()
            U, infos = m.solve_many(F)
            batched = time.perf_counter() - t0
            seq = min(args.seq, len(F))
            t0 = time.perf_counter()
            for k in range(seq):
                u1, _, _ = m.solve(F[k], reactions=False)
                err_seq = float(np.linalg.norm(U[k] - u1) / np.linalg.norm(u1))
            sequential = (time.perf_counter() - t0) / seq
            if U64 is None:
                U64 = U
            err64 = float(max(np.linalg.norm(u - r) / np.linalg.norm(r) for u, r in zip(U, U64)))
            modes[name] = {"batched_s": round(batched, 3), "cases_per_s": round(len(F) / batched, 2),
                           "sequential_s_per_case": round(sequential, 4),
                           "speedup_vs_sequential": round(sequential * len(F) / batched, 2),
                           "iterations_max": max(i["iterations"] for i in infos),
                           "sweeps": m.sweeps(), "sweep_weight": m.sweep_weight(),
                           "waves_per_cu": chosen[0],
                           "cases_per_pass": chosen[1],
                           "rows_per_wave": 64 // max(m.batch_lanes(), 1) if m.batch_lanes() else None,
                           "rel_err_last_checked_vs_sequential": err_seq,
                           "max_rel_err_vs_fp64": err64,
                           "device_retries": sum(i["device_retries"] for i in infos)}
            log(f"beam {name}: {batched:.3f} s for {len(F)} cases ({len(F) / batched:.1f} cases/s), "
                f"sequential {sequential:.4f} s per case, max rel err vs FP64 {err64:.1e}")
    head = min((k for k in modes if modes[k]["max_rel_err_vs_fp64"] <= 1e-8),
               key=lambda k: modes[k]["batched_s"])
    cpu = None
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        s_ = oracle.System(p.nodes, p.conn, p.young, p.nu, p.fixed)
        assert s_.factorize() == oracle.OK
        t_fac = time.perf_counter() - t0
        t0 = time.perf_counter()
        nsamp = min(4, len(F))
        for k in range(nsamp):
            s_.solve(F[k])
        per = (time.perf_counter() - t0) / nsamp
        cpu = {"value": round(len(F) / (t_fac + per * len(F)), 3), "unit": "cases/s",
               "cores": 1, "kind": "port",
               "sample": f"the oracle's FreeSystem build with IC(0) ({t_fac:.2f} s) and "
                         f"{nsamp} of the {len(F)} FP64 PCG solves ({per:.2f} s each) measured, "
                         f"scaled to all of them"}
    h = modes[head]
    emit({"metric": "load cases per second, Tet10 beam 40x8x8 (69,360 free DOF), the faster of "
                    "FP64 and mixed_jacobi within 1e-8 of FP64",
          "value": h["cases_per_s"], "unit": "cases/s", "n_gpus": 1,
          "higher_is_better": True, "data": "synthetic",
          "dtype": "fp64" if head == "fp64" else "fp64+fp32-ic0-jacobi",
          "config": {"workload": "beam_load_cases", "free_dof": p.nfree, "cases": len(F),
                     "precision": head, "sweeps": h["sweeps"], "waves_per_cu": h["waves_per_cu"],
                     "cases_per_pass": h["cases_per_pass"], "rows_per_wave": h["rows_per_wave"]},
          "batched_s": h["batched_s"], "sequential_s_per_case": h["sequential_s_per_case"],
          "speedup_vs_sequential": h["speedup_vs_sequential"], "modes": modes,
          "cpu_baseline": cpu,
          "speedup_vs_cpu": round(h["cases_per_s"] / cpu["value"], 1) if cpu else None})


def run_box_batch(args, ma, meshes):
    """64 Tet4 boxes of uneven size in one batch (ml_batch) against per-model solves."""
    boxes = meshes.box_batch(count=args.models)
    models = [ma.Model(b) for b in boxes]
    log(f"boxes: {len(boxes)} models, {sum(b.nfree for b in boxes)} free DOF in all")
    t0 = time.perf_counter()
    refs = []
    for m in models:
        m.factor()
        u, _, _ = m.solve(reactions=False)
        refs.append(u)
    sequential = time.perf_counter() - t0
    out = {}
    for name, code in (("fp64", ma.FP64), ("mixed_ic0", ma.MIXED_IC0),
                       ("mixed_jacobi", ma.MIXED_JACOBI)):
        t0 = time.perf_counter()
        batch = ma.Batch(models, precision=code)
        setup = time.perf_counter() - t0
        batch.solve()
        t0 = time.perf_counter()
        us, infos = batch.solve()
        solve = time.perf_counter() - t0
        err = max(float(np.linalg.norm(u - r) / np.linalg.norm(r)) for u, r in zip(us, refs))
        out[name] = {"batch_setup_s": round(setup, 3), "batch_solve_s": round(solve, 4),
                     "max_rel_err_vs_separate_fp64": err,
                     "iterations_max": max(i["iterations"] for i in infos)}
        del batch
    cpu = None
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        for b in boxes:
            oracle.model_solve(b, iterative=True)
        wall = time.perf_counter() - t0
        cpu = {"value": round(len(boxes) / wall, 2), "unit": "models/s", "cores": 1,
               "kind": "port", "sample": f"the oracle's whole Model::solve of all {len(boxes)} "
                                         f"models one after another ({wall:.2f} s), measured"}
    valid = [k for k in out if out[k]["max_rel_err_vs_separate_fp64"] <= 1e-8]
    head = min(valid, key=lambda k: out[k]["batch_solve_s"])
    emit({"metric": "models per second, 64 uneven Tet4 boxes (500-20,000 DOF), batched solve",
          "value": round(len(models) / out[head]["batch_solve_s"], 2), "unit": "models/s",
          "n_gpus": 1, "higher_is_better": True, "data": "synthetic", "dtype": head,
          "config": {"workload": "box_batch", "models": len(models),
                     "free_dof": sum(b.nfree for b in boxes)},
          "sequential_factor_and_solve_s": round(sequential, 3), "modes": out,
          "cpu_baseline": cpu,
          "speedup_vs_cpu": (round(len(models) / out[head]["batch_solve_s"] / cpu["value"], 1)
                             if cpu else None)})


def run_tet10(args, ma, meshes):
    """The Tet10 cantilever (60x12x12 cells, 225,000 free DOF): every mode's step (assemble +
    IC(0) + PCG + reactions), iterations and error against FP64 and against the oracle's FP64
    skyline solution (the committed fixture, when present)."""
    p = meshes.cantilever_tet10()
    sky_path = os.path.join(ROOT, "tests", "golden", "cantilever_tet10_60x12x12_skyline.npz")
    u_sky = np.load(sky_path)["displacements"] if os.path.exists(sky_path) else None
    res = {}
    u64 = None
    codes = (("fp64", ma.FP64), ("mixed_ic0", ma.MIXED_IC0), ("mixed_jacobi", ma.MIXED_JACOBI),
             ("mixed", ma.MIXED))
    if args.precision != "best":
        codes = tuple(c for c in codes if c[0] in ("fp64", args.precision))
    for name, code in codes:
        with ma.Model(p, precision=code) as m:
            m.autotune()
            m.factor()
            m.solve(reactions=False)
            steps = []
            for _ in range(max(1, args.steps)):
                t0 = time.perf_counter()
                m.assemble()
                m.factor()
                u, r, info = m.solve()
                steps.append(time.perf_counter() - t0)
            step = min(steps)
            sweeps = m.sweeps()
            weight = m.sweep_weight()
            f = m.loads()
        if u64 is None:
            u64 = u
        balance = float(np.abs(r.reshape(-1, 3).sum(axis=0) + f.reshape(-1, 3).sum(axis=0)).max()
                        / np.abs(f.reshape(-1, 3).sum(axis=0)).max())
        res[name] = {"s_per_step": round(step, 4), "iterations": info["iterations"],
                     "sweeps": sweeps, "sweep_weight": weight,
                     "true_residual": info["true_residual"],
                     "escalated": info["escalated"], "device_retries": info["device_retries"],
                     "rel_err_vs_fp64": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
                     "relative_gap_vs_skyline": (float(np.abs(u - u_sky).max() /
                                                       np.abs(u_sky).max())
                                                 if u_sky is not None else None),
                     "reaction_balance": balance}
        log(f"tet10 {name}: {res[name]}")
    best = min((k for k in res if res[k]["rel_err_vs_fp64"] <= 1e-8), key=lambda k: res[k]["s_per_step"])
    cpu = None
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        ref = oracle.model_solve(p, iterative=True)
        wall = time.perf_counter() - t0
        cpu = {"value": round(wall, 2), "unit": "s", "cores": 1, "kind": "port",
               "sample": f"the oracle's whole Model::solve (FreeSystem build, IC(0), FP64 PCG to "
                         f"1e-12 in {ref['iterations']} iterations, reactions, stresses), "
                         f"measured in full",
               "rel_err_vs_device_fp64": float(np.linalg.norm(ref["displacements"] - u64) /
                                               np.linalg.norm(u64))}
    emit({"metric": "Tet10 cantilever solve s (assemble + IC(0) + PCG), FP64 vs mixed",
          "value": res[best]["s_per_step"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
          "data": "synthetic", "dtype": best,
          "config": {"workload": "cantilever_tet10_60x12x12", "free_dof": p.nfree},
          "modes": res, "cpu_baseline": cpu,
          "speedup_vs_cpu": round(cpu["value"] / res[best]["s_per_step"], 1) if cpu else None})


def run_simp(args, ma, meshes):
    """Adjoint compliance gradient on the SIMP box (80x40x40 cells, 403,440 free DOF): FP64
    solve, device gradient, the oracle's FP64 CPU adjoint timed beside it, and central
    differences on random elements computed without cancellation: with K+- = K +- h k_e,
    (C(E_e + h) - C(E_e - h)) / 2h = -u+.k_e u- exactly, from two perturbed solves (run in
    the Jacobi mode)."""
    import oracle
    p = meshes.simp_box()
    m = ma.Model(p)
    m.factor()
    t0 = time.perf_counter()
    u, _, info = m.solve(reactions=False)
    t_solve = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = m.compliance_gradient()
    t_grad = time.perf_counter() - t0
    rng = np.random.default_rng(6)
    elems = rng.choice(len(p.conn), args.fd, replace=False)
    probe = ma.Model(p, precision=ma.MIXED_JACOBI)
    probe.autotune()
    worst = 0.0
    t0 = time.perf_counter()
    for e in elems:
        h = 1e-4 * p.young[e]
        sols = []
        for sgn in (1.0, -1.0):
            young = p.young.copy()
            young[e] += sgn * h
            probe.set_young(young)
            probe.assemble()
            probe.factor()
            ue, _, _ = probe.solve(reactions=False)
            sols.append(ue)
        dofs = (3 * p.conn[e][:, None] + np.arange(3)).ravel()
        ke = oracle.solid_stiffness(p.nodes[p.conn[e]], 1.0, p.nu[e])
        fd = -float(sols[0][dofs] @ ke @ sols[1][dofs])
        worst = max(worst, abs(fd - g[e]) / abs(g[e]))
    t_fd = time.perf_counter() - t0
    probe.close()
    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        g_cpu = oracle.compliance_gradient(p.nodes, p.conn, p.young, p.nu, u)
        cpu = {"value": round(time.perf_counter() - t0, 3), "unit": "s", "cores": 1,
               "kind": "port", "sample": "the oracle's FP64 CPU adjoint over all elements at the "
                                         "device solution",
               "bit_identical": bool(np.array_equal(g, g_cpu))}
    m.close()
    emit({"metric": "adjoint compliance gradient s, SIMP Tet4 box 80x40x40 (403,440 free DOF)",
          "value": round(t_grad, 4), "unit": "s", "n_gpus": 1, "higher_is_better": False,
          "data": "synthetic", "dtype": "fp64",
          "config": {"workload": "simp_box_80x40x40", "free_dof": p.nfree,
                     "elements": len(p.conn)},
          "solve_s": round(t_solve, 3), "iterations": info["iterations"],
          "fd_elements": args.fd, "fd_relative_step": 1e-4, "fd_max_rel_diff": worst,
          "fd_s": round(t_fd, 2), "cpu_baseline": cpu})


def run_stresses(args, ma, meshes):
    """Post-processing of the cube's solution (jittered cube, 6M Tet4 at 100^3): Gauss-point
    stresses and their average at the nodes, timed with device events; the oracle on one core
    beside it, over the whole mesh."""
    p = meshes.jittered_cube(args.cells)
    m = ma.Model(p)
    m.factor()
    u, _, info = m.solve(reactions=False)
    out = m.nodal_stresses(u)  # the first call builds the node incidence list
    m.profile(enable=1)
    reps = 5
    for _ in range(reps):
        out = m.nodal_stresses(u)
    pr = m.profile(enable=0)
    st_ms = pr["stresses"][0] / pr["stresses"][1]
    nd_ms = pr["nodal"][0] / pr["nodal"][1]
    ne, nen = p.conn.shape
    nn = len(p.nodes)
    ngp = 1 if nen == 4 else 4
    # algorithmic bytes, each array once: Gauss-point stresses read connectivity (int32), the
    # element's coordinates and displacements, E and nu, and write 9 words per Gauss point; the
    # averaging reads the six stress words of each Gauss point, the incidence list, the node
    # positions, and writes 9 words per node
    st_bytes = ne * (4 * nen + 16 + ngp * 72) + nn * (24 + 24)
    nd_bytes = ne * ngp * 48 + 8 * (nn + 1) + 4 * ne * nen + nn * (24 + 72)
    cpu = None
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        gp = oracle.stresses(p.nodes, p.conn, p.young, p.nu, u)
        ref = oracle.nodal_stresses(p.nodes, p.conn, gp)
        cpu = {"value": round(time.perf_counter() - t0, 3), "unit": "s", "cores": 1,
               "kind": "port", "sample": "the oracle's Gauss-point stresses and nodal averaging "
                                         "over the whole mesh at the device solution"}
        cpu["bit_exact"] = bool(np.array_equal(out, ref))
    total = st_ms + nd_ms
    emit({"metric": "Gauss-point + nodal stresses ms, jittered cube (Tet4)",
          "value": round(total, 4), "unit": "ms", "n_gpus": 1, "higher_is_better": False,
          "data": "synthetic", "dtype": "fp64",
          "config": {"workload": f"jittered_cube_{args.cells}_stresses", "elements": int(ne),
                     "nodes": int(nn)},
          "kernels": {"stresses": {"avg_ms": round(st_ms, 4), "bytes": st_bytes,
                                   "achieved_GBs": round(st_bytes / (st_ms * 1e-3) / 1e9, 1)},
                      "nodal": {"avg_ms": round(nd_ms, 4), "bytes": nd_bytes,
                                "achieved_GBs": round(nd_bytes / (nd_ms * 1e-3) / 1e9, 1)}},
          "solve_iterations": info["iterations"], "cpu_baseline": cpu,
          "speedup_vs_cpu": round(cpu["value"] * 1e3 / total, 1) if cpu else None})


def run_plastic(args, ma, meshes):
    """Quasi-static von Mises plasticity (Model::solve_quasi_static) on a Tet4 bar 120x24x24
    (414,720 elements, 224,375 free DOF) bent by a ramped tip displacement over 5 increments:
    every Newton iteration's element responses, tangent assembly, IC(0) and PCG on the device.
    The oracle's Newton (same algorithm, one core) runs the first increment, and its measured
    costs per Newton iteration (responses, FreeSystem build with IC(0)) and per PCG iteration
    are scaled to the device run's counts."""
    from maille_amd import plasticity as P
    p = meshes.plastic_bar()
    codes = {"fp64": ma.FP64, "mixed": ma.MIXED, "mixed_ic0": ma.MIXED_IC0,
             "mixed_jacobi": ma.MIXED_JACOBI}
    prec = "fp64" if args.precision == "best" else args.precision
    m = ma.Model(p, precision=codes[prec])
    if args.sweeps is not None and prec != "fp64":
        m.set_sweeps(args.sweeps)
    elif prec == "mixed_jacobi":
        m.autotune()  # the sweep count whose whole solve of the elastic operator is fastest
    sy, hd = 250.0, 2000.0
    step, amp = P.Step(1.0, args.increments), P.Amplitude.ramp(1.0)
    incs, _ = P.solve_quasi_static(m, P.Step(1.0, 1), P.Amplitude([(0.0, 0.0), (1.0, 1e-3)]),
                                   sy, hd, stresses=False)  # warm-up: an elastic increment
    m.profile(enable=1)
    t0 = time.perf_counter()
    incs, info = P.solve_quasi_static(m, step, amp, sy, hd, stresses=False)
    wall = time.perf_counter() - t0
    pr = m.profile(enable=0)
    ne, nen = p.conn.shape
    ngp = 1 if nen == 4 else 4
    mm = 3 * nen
    # k_response's algorithmic bytes per element, each array once: connectivity (int32),
    # coordinates and displacements of its nodes, E, nu, yield stress, hardening, the committed
    # states (7 words per Gauss point) in; the tangent (m^2), internal force (m), trial states
    # and Gauss-point records (9 words per point) out
    resp_bytes = ne * (4 * nen + 48 * nen + 32 + 56 * ngp + 8 * mm * mm + 8 * mm + 56 * ngp
                       + 72 * ngp)
    resp_ms = pr["stiffness"][0] / max(pr["stiffness"][1], 1)
    tri = pr["precond64"] if prec == "fp64" else pr["precond32"]
    kernels = {
        "response": {"avg_ms": round(resp_ms, 4), "launches": pr["stiffness"][1],
                     "bytes_per_launch": resp_bytes,
                     "achieved_GBs": round(resp_bytes / (resp_ms * 1e-3) / 1e9, 1)},
        "assembly": {"avg_ms": round(pr["assembly"][0] / max(pr["assembly"][1], 1), 4)},
        "ic0": {"avg_ms": round(pr["ic0"][0] / max(pr["ic0"][1], 1), 3), "launches": pr["ic0"][1]},
        "precondition": {"avg_ms": round(tri[0] / max(tri[1], 1), 4), "launches": tri[1],
                         "share_of_run": round(tri[0] * 1e-3 / wall, 3)},
        "spmv": {"avg_ms": round((pr["spmv64"][0] + pr["spmv32"][0]) /
                                 max(pr["spmv64"][1] + pr["spmv32"][1], 1), 4)},
    }
    cpu = None
    if not args.no_cpu:
        import oracle
        first = oracle.quasi_static(p.nodes, p.conn, p.young, p.nu, sy, hd, p.fixed, p.imposed,
                                    p.f, step.time_step, 1, [(0.0, 0.0), (1.0, 1.0)],
                                    iterative=True)
        t_resp, t_build, t_solve = first["times"]
        nwt = max(first["newton"], 1)
        per_newton = (t_resp + t_build) / nwt
        per_pcg = t_solve / max(first["pcg"], 1)
        gap = float(np.linalg.norm(incs[1].displacements - first["u"][1]) /
                    np.linalg.norm(first["u"][1]))
        if m.sweeps() > 0:
            # the device's PCG counts are those of the Jacobi-swept preconditioner, not the
            # oracle's exact IC(0) solves: scale the oracle per Newton iteration instead (its
            # first increment's PCG work per Newton iteration, a lower bound for the plastic
            # increments after it)
            est = (t_resp + t_build + t_solve) / nwt * info["newton_iterations"]
            how = "scaled per Newton iteration to the device run's"
        else:
            est = per_newton * info["newton_iterations"] + per_pcg * info["pcg_iterations"]
            how = "scaled to the device run's counts"
        cpu = {"value": round(est, 2), "unit": "s", "cores": 1, "kind": "port",
               "sample": (f"oracle Newton on the first increment ({first['newton']} Newton, "
                          f"{first['pcg']} PCG iterations measured: {per_newton:.2f} s per Newton "
                          f"iteration for responses, FreeSystem build and IC(0), "
                          f"{per_pcg * 1e3:.1f} ms per PCG iteration), {how} "
                          f"{info['newton_iterations']} Newton iterations"),
               "first_increment_rel_l2": gap}
    emit({"metric": "quasi-static von Mises Newton-Raphson s, Tet4 bar 120x24x24, 5 increments",
          "value": round(wall, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
          "data": "synthetic", "dtype": "fp64" if prec == "fp64" else f"fp64+{prec}",
          "config": {"workload": "plastic_bar_120x24x24", "free_dof": p.nfree,
                     "elements": int(ne), "increments": args.increments,
                     "yield_stress": sy, "hardening": hd, "precision": prec,
                     "sweeps": m.sweeps(), "sweep_weight": m.sweep_weight()},
          "newton_iterations": info["newton_iterations"], "cuts": info["cuts"],
          "pcg_iterations": info["pcg_iterations"],
          "residuals": [[float(f"{r:.3e}") for r in inc.residuals] for inc in incs[1:]],
          "max_p": float(incs[-1].equivalent_plastic_strains.max()),
          "phases_s": {"responses": round(info["seconds"][0], 3),
                       "assembly_ic0": round(info["seconds"][1], 3),
                       "pcg": round(info["seconds"][2], 3)},
          "kernels": kernels, "cpu_baseline": cpu,
          "speedup_vs_cpu": round(cpu["value"] / wall, 1) if cpu else None})


def run_transient(args, ma, meshes):
    """Transient heat conduction (Model::solve_step, implicit Euler) on a Tet4 cube of
    60^3 cells (1,296,000 elements, 226,981 nodes): T = 100 on x = 0, convection (h = 10, sink 0)
    on x = 1, from T = 0, 50 steps of 0.002 (k = 1, rho c = 1). The left matrix is assembled and
    factored once; each step is a device PCG solve from zero, as Pcg::solve starts. The oracle
    (one core) runs the same model for 3 steps; its build and per-step costs are scaled to 50."""
    from maille_amd import thermal as T
    n = args.cells if args.cells != 100 else 60
    nodes, conn = meshes.box_grid((n, n, n), (1.0, 1.0, 1.0))
    nn, ne = len(nodes), len(conn)
    fixed = (np.abs(nodes[:, 0]) < 1e-12).astype(np.uint8)
    imposed = np.where(fixed, 100.0, 0.0)
    faces = ((0, 1, 2), (0, 3, 1), (1, 3, 2), (2, 3, 0))
    on = np.abs(nodes[:, 0] - 1.0) < 1e-12
    conv = [(e, fc, 10.0, 0.0) for e in np.nonzero(on[conn].sum(axis=1) >= 3)[0]
            for fc, vs in enumerate(faces) if on[conn[e, list(vs)]].all()]
    steps, period = args.increments if args.increments != 5 else 50, 0.1
    codes = {"fp64": ma.FP64, "mixed": ma.MIXED, "mixed_ic0": ma.MIXED_IC0,
             "mixed_jacobi": ma.MIXED_JACOBI}
    prec = "fp64" if args.precision == "best" else args.precision
    t0 = time.perf_counter()
    m = T.ThermalModel(nodes, conn, 1.0, 1.0, fixed, imposed, convection=conv,
                       precision=codes[prec])
    t_create = time.perf_counter() - t0
    zero = np.zeros(nn)
    m.solve_step(T.Step(period / steps, 1), zero, fluxes=False)  # warm-up: the same dt, factored
    t0 = time.perf_counter()
    incs, info = m.solve_step(T.Step(period, steps), zero, fluxes=False)
    wall = time.perf_counter() - t0
    cpu = None
    if not args.no_cpu:
        import oracle
        f = oracle.thermal_loads(nodes, conn, zero, (), conv)
        sample = 3
        run = oracle.transient(nodes, conn, 1.0, 1.0, fixed, imposed, f, zero, period / steps * sample,
                               sample, convection=conv)
        t_build, t_steps = run["times"]
        est = t_build + t_steps / sample * steps
        gap = float(np.abs(incs[sample].temperatures - run["temperatures"][sample]).max() / 100.0)
        cpu = {"value": round(est, 2), "unit": "s", "cores": 1, "kind": "port",
               "sample": (f"oracle on the same model: FreeSystem build with IC(0) {t_build:.2f} s "
                          f"and {sample} steps ({t_steps / sample:.2f} s each, {run['pcg']} PCG "
                          f"iterations) measured, scaled to {steps} steps"),
               "step3_max_abs_diff_over_100": gap}
    emit({"metric": "transient heat conduction s, Tet4 cube 60^3, 50 implicit Euler steps",
          "value": round(wall, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
          "data": "synthetic", "dtype": "fp64" if prec == "fp64" else f"fp64+{prec}",
          "config": {"workload": f"thermal_cube_{n}", "nodes": int(nn), "elements": int(ne),
                     "steps": steps, "convection_faces": len(conv)},
          "ms_per_step": round(wall * 1e3 / steps, 3), "pcg_iterations": info["iterations"],
          "create_s": round(t_create, 2), "factor_s": round(info["seconds"][1], 3),
          "cpu_baseline": cpu, "speedup_vs_cpu": round(cpu["value"] / wall, 1) if cpu else None})


def run_large(args, ma, meshes):
    """Large strains (Kinematics::Large) with follower pressures: a Neo-Hookean Tet4 block of
    40x40x40 cells (384,000 elements, 203,043 free DOF), symmetry supports on x = 0, y = 0,
    z = 0 and a follower pressure on its three other faces ramped to the value that compresses
    it to a stretch of 0.9, in 4 increments; every Newton iteration's element responses,
    follower terms, assembly, IC(0) and PCG on the device. The oracle's Newton runs the first
    increment; its costs per Newton and per PCG iteration are scaled to the device run."""
    from maille_amd import plasticity as P
    n = 40
    size = 1.0
    nodes, conn = meshes.box_grid((n, n, n), (size, size, size))
    ne = len(conn)
    young, nu = 210000.0, 0.3
    lam = young * nu / ((1 + nu) * (1 - 2 * nu))
    mu = young / (2 * (1 + nu))
    stretch = 0.9
    jj = stretch ** 3
    p = -(mu * (stretch * stretch - 1.0) + lam * np.log(jj)) / jj
    fixed = np.zeros(3 * len(nodes), dtype=bool)
    for c in range(3):
        fixed[3 * np.nonzero(np.abs(nodes[:, c]) < 1e-12)[0] + c] = True
    faces = ((0, 1, 2), (0, 3, 1), (1, 3, 2), (2, 3, 0))
    pressures = []
    for c in range(3):
        on = np.abs(nodes[:, c] - size) < 1e-12
        for e in np.nonzero(on[conn].sum(axis=1) >= 3)[0]:
            for fc, vs in enumerate(faces):
                if on[conn[e, list(vs)]].all():
                    pressures.append((int(e), fc, p))
    prob = meshes.Problem("block", nodes, conn, np.full(ne, young), np.full(ne, nu), fixed,
                          np.zeros(3 * len(nodes)), np.zeros(3 * len(nodes)))
    codes = {"fp64": ma.FP64, "mixed": ma.MIXED, "mixed_ic0": ma.MIXED_IC0,
             "mixed_jacobi": ma.MIXED_JACOBI}
    prec = "fp64" if args.precision == "best" else args.precision
    m = ma.Model(prob, precision=codes[prec])
    if args.sweeps is not None and prec != "fp64":
        m.set_sweeps(args.sweeps)
    elif prec == "mixed_jacobi":
        m.autotune()  # the sweep count whose whole solve of the initial operator is fastest
    step, amp = P.Step(1.0, 4), P.Amplitude.ramp(1.0)
    P.solve_large_strain(m, P.Step(1.0, 1), P.Amplitude([(0.0, 0.0), (1.0, 1e-3)]),
                         neo_hookean=True, pressures=pressures)  # warm-up
    m.profile(enable=1)
    t0 = time.perf_counter()
    incs, info = P.solve_large_strain(m, step, amp, neo_hookean=True, pressures=pressures)
    wall = time.perf_counter() - t0
    pr = m.profile(enable=0)
    resp_ms = pr["stiffness"][0] / max(pr["stiffness"][1], 1)
    tri = pr["precond64"]
    u = incs[-1].displacements.reshape(-1, 3)
    hom = float(np.abs(u - (stretch - 1.0) * nodes).max())
    cpu = None
    if not args.no_cpu:
        import oracle
        t1 = time.perf_counter()
        # the device run's first increment: lambda = 0.25 at t = 0.25
        first = oracle.large_strain(nodes, conn, young, nu, fixed, np.zeros(3 * len(nodes)),
                                    np.zeros(3 * len(nodes)), step.time_step, 1,
                                    [(0.0, 0.0), (step.time_step, 0.25)], neo_hookean=True,
                                    pressures=pressures)
        t_first = time.perf_counter() - t1
        per_newton = t_first / max(first["newton"], 1)
        est = per_newton * info["newton_iterations"]
        cpu = {"value": round(est, 2), "unit": "s", "cores": 1, "kind": "port",
               "sample": (f"oracle Newton on the first increment ({first['newton']} Newton, "
                          f"{first['pcg']} PCG iterations, {t_first:.1f} s), scaled per Newton "
                          f"iteration to the device run's {info['newton_iterations']}"),
               "first_increment_rel_l2": float(np.linalg.norm(incs[1].displacements - first["u"][1])
                                               / np.linalg.norm(first["u"][1]))}
    emit({"metric": "large-strain Neo-Hookean Newton-Raphson with follower pressures s, Tet4 "
                    "block 40^3, 4 increments",
          "value": round(wall, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
          "data": "synthetic", "dtype": "fp64" if prec == "fp64" else f"fp64+{prec}",
          "config": {"workload": "neo_hookean_block_40", "elements": int(ne),
                     "free_dof": prob.nfree, "follower_faces": len(pressures),
                     "pressure": round(p, 3), "stretch": stretch, "precision": prec,
                     "sweeps": m.sweeps(), "sweep_weight": m.sweep_weight()},
          "newton_iterations": info["newton_iterations"], "cuts": info["cuts"],
          "pcg_iterations": info["pcg_iterations"],
          "residuals": [[float(f"{r:.3e}") for r in inc.residuals] for inc in incs[1:]],
          "max_dev_from_homogeneous": hom,
          "phases_s": {"responses_follower": round(info["seconds"][0], 3),
                       "assembly_ic0": round(info["seconds"][1], 3),
                       "pcg": round(info["seconds"][2], 3)},
          "kernels": {"large_response": {"avg_ms": round(resp_ms, 4),
                                         "launches": pr["stiffness"][1]},
                      "precondition": {"avg_ms": round(tri[0] / max(tri[1], 1), 4),
                                       "launches": tri[1],
                                       "share_of_run": round(tri[0] * 1e-3 / wall, 3)}},
          "cpu_baseline": cpu, "speedup_vs_cpu": round(cpu["value"] / wall, 1) if cpu else None})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=100)
    ap.add_argument("--precision", choices=["best", "mixed", "mixed_ic0", "mixed_jacobi", "fp64"],
                    default="best",
                    help="mode(s) to time; 'best' compares fp64, mixed_ic0 and mixed_jacobi and "
                         "reports the fastest whose displacements stay within 1e-8 of the FP64 "
                         "drop-in")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="end with a plain 256 MiB read + 256 MiB write kernel, the reference "
                         "against which HBM counters (FETCH_SIZE, WRITE_SIZE) are scaled")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="PCG iterations of the CPU sample (0: a tenth of the FP64 count)")
    ap.add_argument("--sweeps", type=int, default=None,
                    help="Jacobi sweeps of the mixed modes (default: the autotuner's choice)")
    ap.add_argument("--compare-steps", type=int, default=2,
                    help="steps timed per mode to pick the headline mode")
    ap.add_argument("--workload", choices=["cube", "beam_cases", "box_batch", "tet10", "simp",
                                           "stresses", "plastic", "transient", "large"],
                    default="cube", help="cube: the headline 3.1M-DOF solve; the others are the "
                    "batched, Tet10 and adjoint workloads (one JSON line each)")
    ap.add_argument("--cases", type=int, default=256)
    ap.add_argument("--seq", type=int, default=8, help="sequential solves timed (beam_cases)")
    ap.add_argument("--models", type=int, default=64)
    ap.add_argument("--fd", type=int, default=32, help="elements checked by central differences")
    ap.add_argument("--increments", type=int, default=5, help="Step increments (plastic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding "
              "of this path is a design, not a build", file=sys.stderr)
        sys.exit(2)

    import torch

    import maille_amd as ma
    from maille_amd import meshes

    if args.workload != "cube":
        {"beam_cases": run_beam_cases, "box_batch": run_box_batch, "tet10": run_tet10,
         "simp": run_simp, "stresses": run_stresses,
         "plastic": run_plastic, "transient": run_transient,
         "large": run_large}[args.workload](args, ma, meshes)
        return

    t0 = time.time()
    problem = meshes.jittered_cube(args.cells)
    log(f"mesh: {len(problem.conn)} elements, {problem.nfree} free DOF ({time.time() - t0:.1f} s)")

    codes = {"fp64": ma.FP64, "mixed": ma.MIXED, "mixed_ic0": ma.MIXED_IC0,
             "mixed_jacobi": ma.MIXED_JACOBI}
    modes = ["fp64", "mixed_ic0", "mixed_jacobi"] if args.precision == "best" else \
        (["fp64"] if args.precision == "fp64" else ["fp64", args.precision])
    models, create = {}, {}
    for prec in modes:
        t1 = time.perf_counter()
        models[prec] = ma.Model(problem, precision=codes[prec])
        t_create = time.perf_counter() - t1
        if args.sweeps is not None and prec != "fp64":
            models[prec].set_sweeps(args.sweeps)
        chosen = models[prec].autotune()
        create[prec] = {"create_s": round(t_create, 3),
                        "host_pattern_s": round(models[prec].create_info["seconds"][5], 3),
                        "autotune_s": round(time.perf_counter() - t1 - t_create, 3),
                        "waves_per_cu": chosen[0], "reported_occupancy": chosen[2],
                        "sweeps": models[prec].sweeps(),
                        "sweep_weight": models[prec].sweep_weight(),
                        "residency_reported_used": models[prec].residency()}
        log(f"{prec}: model created in {t_create:.1f} s (host pattern "
            f"{create[prec]['host_pattern_s']:.1f} s), autotuned in "
            f"{create[prec]['autotune_s']:.1f} s: {chosen[0]} waves per CU of {chosen[2]}, "
            f"{create[prec]['sweeps']} Jacobi sweeps")
    m0 = models["fp64"]
    f = m0.loads()
    dev = torch.device("cuda:0")
    f_dev = torch.from_numpy(f).to(dev)
    u_dev = {k: torch.zeros(problem.ndof, dtype=torch.float64, device=dev) for k in models}
    r_dev = torch.zeros(problem.ndof, dtype=torch.float64, device=dev)

    def step(prec):
        m = models[prec]
        m.assemble()
        m.factor()
        return m.solve_device(f_dev.data_ptr(), u_dev[prec].data_ptr(), r_dev.data_ptr())

    def timed(prec, warmup, steps):
        m = models[prec]
        retries = 0
        for _ in range(warmup):
            retries += step(prec)["device_retries"]
        torch.cuda.synchronize()
        m.profile(enable=1)
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(steps):
            info = step(prec)
            retries += info["device_retries"]
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t_start
        return {"s": elapsed / steps, "elapsed": elapsed, "info": info,
                "prof": m.profile(enable=0), "device_retries": retries, "steps": steps}

    # every mode is compared on a few steps; the fastest whose displacements stay within 1e-8
    # of the FP64 drop-in is then timed over exactly --steps steps
    compare = {}
    for prec in modes:
        compare[prec] = timed(prec, 1, args.compare_steps)
        log(f"{prec}: {compare[prec]['s']:.4f} s per step over {args.compare_steps}, "
            f"{compare[prec]['info']['iterations']} iterations, true residual "
            f"{compare[prec]['info']['true_residual']:.2e}, retries "
            f"{compare[prec]['device_retries']}")
    u64 = u_dev["fp64"].cpu().numpy()
    for prec in modes:
        u = u_dev[prec].cpu().numpy()
        compare[prec]["rel_err_vs_fp64"] = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
    valid = [p for p in modes if compare[p]["rel_err_vs_fp64"] <= 1e-8]
    head = min(valid, key=lambda p: compare[p]["s"])
    res = timed(head, args.warmup, args.steps)
    log(f"headline {head}: {res['s']:.4f} s per step over {args.steps} steps; kernels "
        f"{res['prof']}")
    u = u_dev[head].cpu().numpy()
    res["rel_err_vs_fp64"] = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
    per_step, elapsed, info, prof = res["s"], res["elapsed"], res["info"], res["prof"]

    n, nnz_l = m0.nfree(), m0.nnz()
    nnz_full = 2 * nnz_l - n
    sweeps = models[head].sweeps()
    aval = 4 if head == "mixed" else 8
    mval = 8 if head == "fp64" else 4
    kinds = {"spmv": "spmv32" if head == "mixed" else "spmv64",
             "precondition": "precond64" if head == "fp64" else "precond32"}
    spmv_form, spmv_idx = models[head].spmv_index()
    col_bytes = models[head].index_bytes()[1]
    col_bytes = col_bytes or 4
    fused = models[head].sweeps_fused() if sweeps else 0  # register capacity, 0: per launch
    pre_bytes = (2 * fused_sweep_bytes(n, nnz_l - n, sweeps) if fused else
                 2 * sweeps * sweep_bytes(n, nnz_l - n, col_bytes) if sweeps
                 else 2 * trisolve_bytes(n, nnz_l, mval))
    bytes_per = {"spmv": spmv_bytes(n, nnz_full, aval, spmv_idx), "precondition": pre_bytes}
    kernels = {}
    for name, kind in kinds.items():
        ms, cnt = prof[kind]
        avg = ms / max(cnt, 1)
        achieved = bytes_per[name] / (avg * 1e-3) / 1e9 if cnt else 0.0
        kernels[name] = {"avg_ms": round(avg, 4), "launches": cnt,
                         "bytes_per_launch": bytes_per[name],
                         "achieved_GBs": round(achieved, 1),
                         "share_of_step": round(ms / 1e3 / elapsed, 3)}
    kernels["spmv"]["index_form"] = {1: "one 16-bit column stream per group of rows with "
                                         "equal columns (a node's DOFs)",
                                      2: "16-bit column offsets per entry",
                                      4: "int32 columns"}[spmv_form]
    spmv_traffic = measured_spmv_traffic(f"jittered_cube_{args.cells}", head, spmv_form)
    if spmv_traffic and kernels["spmv"]["launches"]:
        # the same launches' rate on the HBM bytes the committed counter passes measured
        rate = spmv_traffic / (kernels["spmv"]["avg_ms"] * 1e-3) / 1e9
        kernels["spmv"]["measured_traffic"] = {
            "bytes_per_launch": spmv_traffic, "GBs": round(rate, 1),
            "frac_of_peak": round(rate / HBM_PEAK_GBS, 4),
            "source": "replayed from the committed FETCH_SIZE/WRITE_SIZE passes, not measured "
                      "in this run"}
    kernels["precondition"]["form"] = (
        f"{sweeps} Jacobi sweeps per triangular system, all in one launch per system "
        f"(k_sweeps_fused, 2 launches, {fused} coefficients per row in registers)" if fused else
        f"{sweeps} Jacobi sweeps per triangular system (k_sweep, {2 * sweeps} launches, "
        f"{col_bytes}-byte columns)" if sweeps else
        "exact dataflow triangular solves (k_trisolve, 2 launches)")
    ic_ms, ic_cnt = prof["ic0"]
    kernels["ic0_factor"] = {"ms_per_step": round(ic_ms / args.steps, 2),
                             "share_of_step": round(ic_ms / 1e3 / elapsed, 3)}
    dominant = max(("spmv", "precondition"), key=lambda k: kernels[k]["share_of_step"])
    d = kernels[dominant]
    if dominant == "precondition" and fused:
        # the application is two launches of k_sweeps_fused, one per triangular system
        per_ms = d["avg_ms"] / 2
        one = fused_sweep_bytes(n, nnz_l - n, sweeps)
        name, achieved, per_launch = "k_sweeps_fused", one / (per_ms * 1e-3) / 1e9, one
    elif dominant == "precondition" and sweeps:
        # the application is 2 x sweeps launches of one kernel, k_sweep: its roofline is that
        # of one launch (one sweep), timed as the application's events over its launch count
        per_ms = d["avg_ms"] / (2 * sweeps)
        one = sweep_bytes(n, nnz_l - n, col_bytes)
        name, achieved, per_launch = "k_sweep", one / (per_ms * 1e-3) / 1e9, one
    else:
        name, achieved = ("k_spmv" if dominant == "spmv" else "k_trisolve"), d["achieved_GBs"]
        per_ms, per_launch = d["avg_ms"], d["bytes_per_launch"]
    roofline = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(name, f"jittered_cube_{args.cells}", head, sweeps,
                                            fused),
                "avg_launch_ms": round(per_ms, 5), "algorithmic_bytes_per_launch": per_launch}
    if roofline["traffic"] is not None:
        roofline["traffic_source"] = ("replayed from the committed FETCH_SIZE/WRITE_SIZE passes "
                                      f"profiles/traffic_{name}_jittered_cube_{args.cells}_"
                                      f"{head}.json, not measured in this run")
    p64 = compare["fp64"]["prof"]
    spmv64_ms, spmv64_cnt = p64["spmv64"]
    spmv64_avg = spmv64_ms / max(spmv64_cnt, 1)
    spmv64_gbs = (spmv_bytes(n, nnz_full, 8, m0.spmv_index()[1]) / (spmv64_avg * 1e-3) / 1e9
                  if spmv64_cnt else 0.0)

    # the whole Model::solve (model.rs:548-610) of the headline mode: model creation (mesh
    # checks, free numbering, Graph, RCM, pattern on the host; upload, element stiffness and
    # assembly on the device), IC(0), PCG, reactions, Gauss-point stresses
    mh = models[head]
    gp = torch.zeros(len(problem.conn) * 9, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mh.assemble()  # element stiffness + assembly (synchronised)
    t2 = time.perf_counter()
    mh.factor()
    t3 = time.perf_counter()
    e2e = mh.solve_device(f_dev.data_ptr(), u_dev[head].data_ptr(), r_dev.data_ptr())
    t4 = time.perf_counter()
    mh.stresses_device(u_dev[head].data_ptr(), gp.data_ptr())
    torch.cuda.synchronize()
    t5 = time.perf_counter()
    del gp
    end_to_end = {
        "create_s": create[head]["create_s"], "host_pattern_s": create[head]["host_pattern_s"],
        "assemble_s": round(t2 - t1, 4), "factor_s": round(t3 - t2, 4),
        "solve_s": round(e2e["seconds"][2], 4), "reactions_s": round(e2e["seconds"][3], 4),
        "stresses_s": round(t5 - t4, 4),
        "note": "create = mesh checks, free numbering, Graph, RCM and pattern on the host, "
                "upload, stiffness and assembly on the device (includes assemble_s once)",
    }
    end_to_end["total_s"] = round(create[head]["create_s"] + (t5 - t2), 3)

    cpu = None
    if not args.no_cpu:
        t_cpu = time.time()
        iters = compare["fp64"]["info"]["iterations"]
        sample = args.cpu_iters or max(8, -(-iters // 10))
        cpu = cpu_baseline(problem, iters, sample)
        log(f"cpu baseline (oracle, 1 core): {cpu['value']} s estimated; "
            f"{cpu['multicore']['cores']} threads: {cpu['multicore']['value']} s "
            f"({time.time() - t_cpu:.1f} s sample)")

    out = {
        "metric": METRIC, "value": round(per_step, 4), "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(per_step * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "fp64", "mixed": "fp32+fp64", "mixed_ic0": "fp64+fp32-ic0",
                  "mixed_jacobi": "fp64+fp32-ic0-jacobi"}[head],
        "data": "synthetic",
        "config": {"workload": f"jittered_cube_{args.cells}", "free_dof": n, "nnz_L": nnz_l,
                   "elements": len(problem.conn), "precision": head, "sweeps": sweeps,
                   "sweep_weight": models[head].sweep_weight(),
                   "step": "assemble + IC(0) + PCG to 1e-12 + reactions, inputs in HBM",
                   "dag_depth": m0.levels()},
        "iterations": int(info["iterations"]), "true_residual": info["true_residual"],
        "rel_err_vs_fp64": res["rel_err_vs_fp64"],
        "device_retries": int(res["device_retries"]),
        "roofline": roofline, "kernels": kernels,
        "spmv64": {"avg_ms": round(spmv64_avg, 4), "achieved_GBs": round(spmv64_gbs, 1),
                   "frac": round(spmv64_gbs / HBM_PEAK_GBS, 4)},
        "modes": {p: {"s_per_step": round(compare[p]["s"], 4),
                      "steps": compare[p]["steps"],
                      "iterations": int(compare[p]["info"]["iterations"]),
                      "sweeps": create[p]["sweeps"],
                      "residual_replacements": int(compare[p]["info"]["refinements"]),
                      "escalated_to_fp64_operator": bool(compare[p]["info"]["escalated"]),
                      "true_residual": compare[p]["info"]["true_residual"],
                      "rel_err_vs_fp64": compare[p]["rel_err_vs_fp64"],
                      "device_retries": int(compare[p]["device_retries"]),
                      "setup": create[p]} for p in modes},
        "end_to_end": end_to_end,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(cpu["value"] / per_step, 1) if cpu else None,
        "speedup_vs_multicore_cpu": round(cpu["multicore"]["value"] / per_step, 1) if cpu else None,
    }
    print(json.dumps(out), flush=True)
    # deterministic teardown: every device handle is freed before the interpreter exits
    for m in models.values():
        m.close()
    del f_dev, u_dev, r_dev
    torch.cuda.synchronize()
    if args.copy_probe:
        copy_probe(torch)


def copy_probe(torch):
    """One elementwise kernel that reads 256 MiB and writes 256 MiB: what HBM counters read for
    it scales their readings for the solve's kernels."""
    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    log("copy probe: 268435456 bytes read, 268435456 bytes written")


def measured_spmv_traffic(workload, precision, index_form):
    """HBM bytes per launch of the FP64 SpMV in the index form in use (1: grouped column
    streams, 2: 16-bit offsets per entry), from the committed counter passes of the fused sweeps
    on this workload and precision (the SpMV's launches of the same runs); None otherwise."""
    path = os.path.join(ROOT, "profiles", f"traffic_k_sweeps_fused_{workload}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    name = {1: "k_spmv_grouped<double, 8>", 2: "k_spmv<double, 8, true>"}.get(index_form)
    other = rec.get("other_kernels_same_runs", {}).get(name)
    if not other or "fetch_bytes" not in other or "write_bytes" not in other:
        return None
    return int(other["fetch_bytes"] + other["write_bytes"])


def measured_traffic(kernel, workload, precision, sweeps, variant):
    """HBM bytes per launch of `kernel` from the committed counter passes (profiles/), when
    they were taken on this workload and precision with the same sweep count and fused-kernel
    variant as this run; None otherwise."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("sweeps", sweeps) != sweeps or rec.get("variant", variant) != variant:
        return None
    return rec.get("bytes_per_launch")


if __name__ == "__main__":
    main()
