"""Pins the oracle (the FP64 CPU restatement of the reference path) without running the
reference: closed forms, invariants, brute-force reductions and the reference's own
analytic tests, restated. Each test names the reference test or property it follows."""
import itertools

import numpy as np
import pytest

import oracle
from maille_amd import meshes as M
from conftest import relative_gap

E, NU = 210e3, 0.3


# ------------------------------------------------------------------ shape and quadrature
@pytest.mark.parametrize("nen", [4, 10])
def test_shape_functions_kronecker_and_partition_of_unity(nen):
    """shape.rs:336-374: N_i(node_j) = delta_ij, sum N = 1, sum dN = 0."""
    ref = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    if nen == 10:
        ref = np.vstack([ref, [(ref[a] + ref[b]) / 2 for a, b in M.TET10_EDGES]])
    for j, p in enumerate(ref):
        n, dn = oracle.tet_shape(nen, *p)
        assert np.allclose(n, np.eye(nen)[j], atol=1e-15)
    n, dn = oracle.tet_shape(nen, 0.2, 0.3, 0.1)
    assert abs(n.sum() - 1) < 1e-15 and np.abs(dn.sum(axis=0)).max() < 1e-14


def test_tet10_derivatives_match_finite_differences_and_reproduce_quadratics():
    """shape.rs:348-374."""
    p = np.array([0.21, 0.17, 0.33])
    n0, dn = oracle.tet_shape(10, *p)
    h = 1e-6
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        fd = (oracle.tet_shape(10, *(p + e))[0] - oracle.tet_shape(10, *(p - e))[0]) / (2 * h)
        assert np.abs(fd - dn[:, c]).max() < 1e-8
    ref = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    nodes = np.vstack([ref, [(ref[a] + ref[b]) / 2 for a, b in M.TET10_EDGES]])
    f = lambda x: 1 + 2 * x[0] - x[1] + 3 * x[0] * x[2] + x[1] ** 2  # noqa: E731
    assert abs(np.dot(n0, [f(x) for x in nodes]) - f(p)) < 1e-14


@pytest.mark.parametrize("nen", [4, 10])
def test_mapping_determinant_is_six_volumes_and_gradient_of_x_is_identity(nen):
    """shape.rs:377-405."""
    rng = np.random.default_rng(1)
    x4 = np.array([[0, 0, 0], [2, 0.1, 0], [0.3, 1.5, 0.2], [0.1, 0.2, 1.2]]) + rng.normal(0, 0.05, (4, 3))
    x = x4 if nen == 4 else np.vstack([x4, [(x4[a] + x4[b]) / 2 for a, b in M.TET10_EDGES]])
    six_v = M.signed_volume(x4, np.array([[0, 1, 2, 3]]))[0]
    for q in [(0.25, 0.25, 0.25), (0.1, 0.2, 0.3)]:
        _, dn = oracle.tet_shape(nen, *q)
        det, g = oracle.mapping3d(x, dn)
        assert abs(det - six_v) < 1e-13 * abs(six_v)
        assert np.abs(x.T @ g - np.eye(3)).max() < 1e-13


# ------------------------------------------------------------------ element
def test_solid_d_is_the_lame_matrix():
    """element.rs:770-779: lambda + 2mu on the diagonal, lambda off it, mu in shear."""
    lam = E * NU / ((1 + NU) * (1 - 2 * NU))
    mu = E / (2 * (1 + NU))
    d = oracle.solid_d(E, NU)
    expect = np.zeros((6, 6))
    expect[:3, :3] = lam
    expect[np.arange(3), np.arange(3)] = lam + 2 * mu
    expect[np.arange(3, 6), np.arange(3, 6)] = mu
    assert np.allclose(d, expect, rtol=1e-14, atol=0)


def _random_tet(nen, seed=3):
    rng = np.random.default_rng(seed)
    x4 = np.array([[0, 0, 0], [1.2, 0.1, 0], [0.2, 0.9, 0.1], [0.1, 0.3, 1.1]]) + rng.normal(0, 0.05, (4, 3))
    if nen == 4:
        return x4
    return np.vstack([x4, [(x4[a] + x4[b]) / 2 for a, b in M.TET10_EDGES]])


@pytest.mark.parametrize("nen", [4, 10])
def test_stiffness_symmetry_rigid_modes_and_energy(nen):
    """element.rs:782-839: K symmetric, the 6 rigid modes in its null space, and for a linear
    field u = eps x the energy u'Ku equals V eps'D eps."""
    x = _random_tet(nen)
    k = oracle.solid_stiffness(x, E, NU)
    assert np.abs(k - k.T).max() < 1e-10 * np.abs(k).max()
    modes = []
    for c in range(3):
        t = np.zeros((len(x), 3))
        t[:, c] = 1
        modes.append(t.ravel())
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1
        modes.append(np.cross(w, x).ravel())
    for mvec in modes:
        assert np.abs(k @ mvec).max() < 1e-9 * np.abs(k).max() * np.abs(mvec).max()
    eps = np.array([1e-3, -2e-3, 0.5e-3, 3e-3, -1e-3, 2e-3])
    grad = np.array([[eps[0], eps[5] / 2, eps[4] / 2],
                     [eps[5] / 2, eps[1], eps[3] / 2],
                     [eps[4] / 2, eps[3] / 2, eps[2]]])
    u = (x @ grad.T).ravel()
    vol = abs(M.signed_volume(x[:4], np.array([[0, 1, 2, 3]]))[0]) / 6
    energy = vol * eps @ oracle.solid_d(E, NU) @ eps
    assert abs(u @ k @ u - energy) < 1e-9 * abs(energy)


@pytest.mark.parametrize("nen", [4, 10])
def test_stresses_of_a_linear_field_are_d_eps(nen):
    """element.rs:840-860."""
    x = _random_tet(nen)
    eps = np.array([1e-3, -2e-3, 0.5e-3, 3e-3, -1e-3, 2e-3])
    grad = np.array([[eps[0], eps[5] / 2, eps[4] / 2],
                     [eps[5] / 2, eps[1], eps[3] / 2],
                     [eps[4] / 2, eps[3] / 2, eps[2]]])
    s = oracle.solid_stresses(x, E, NU, (x @ grad.T).ravel())
    sig = oracle.solid_d(E, NU) @ eps
    assert np.abs(s[:, 3:] - sig).max() < 1e-9 * np.abs(sig).max()


@pytest.mark.parametrize("nen", [4, 10])
def test_face_pressure_splits_the_resultant(nen):
    """assembly.rs:1299-1326: Tet4 face -> -p A n / 3 per vertex; Tet10 face -> 0 at the
    vertices and -p A n / 3 per mid-node; the normal points out of the element."""
    x = _random_tet(nen)
    for face in range(4):
        pos, forces = oracle.face_pressure(x, face, 2.0)
        v = x[pos[:3]]
        cr = np.cross(v[1] - v[0], v[2] - v[0])
        area = np.linalg.norm(cr) / 2
        n = cr / np.linalg.norm(cr)
        if np.dot(n, v.mean(axis=0) - x[:4].mean(axis=0)) < 0:
            n = -n
        per = -2.0 * area * n / 3
        if nen == 4:
            assert np.abs(forces - per).max() < 1e-12
        else:
            assert np.abs(forces[:3]).max() < 1e-12
            assert np.abs(forces[3:] - per).max() < 1e-12


# ------------------------------------------------------------------ dot, graph, RCM
def test_dot_matches_the_plain_sum_for_every_length_around_the_chunk_size():
    """solve.rs:349-359."""
    for n in range(40):
        a = 0.5 + np.arange(n, dtype=float)
        b = 2.0 - 0.25 * np.arange(n, dtype=float)
        plain = float(np.sum(a * b))
        assert abs(oracle.dot(a, b) - plain) <= 1e-12 * (1 + abs(plain))


_B = float(2 ** 53)  # B + 1 rounds to B (ties to even); B + 2 is exact


@pytest.mark.parametrize("a,expected", [
    # lanes: a[0], a[8] accumulate in lane 0 and a[4], a[12] in lane 4, so lane 0 is
    # B - B = 0, lane 4 is 1 + 1 = 2 and the dot is 2. A sequential sum (B, B + 1 = B,
    # 0, 1) and four lanes (lane 0: B + 1 - B + 1 = 1) both give 1.
    ([_B, 0, 0, 0, 1, 0, 0, 0, -_B, 0, 0, 0, 1, 0, 0, 0], 2.0),
    # the combination tree ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)): the left half is
    # B, the right half 1 + 1 = 2, the dot B + 2. Lanes summed in a line (B + 1 = B twice),
    # paired as (s0 + s4) + (s1 + s5) + ... (B + 1 = B), a sequential sum and four lanes
    # (B + 1 = B) all give B.
    ([_B, 0, 0, 0, 1, 1, 0, 0], _B + 2),
    # the tail after the last full chunk is summed on its own, in order from the first
    # element (B, B + 1 = B, B + 1 = B), and added to the chunks' tree last: B. Summed from its
    # end (1 + 1 = 2, 2 + B) it would give B + 2.
    ([0, 0, 0, 0, 0, 0, 0, 0, _B, 1, 1], _B),
    # the tail's own sum (1 + 1 = 2) meets the chunk's B only at the end: B + 2. Folding the
    # tail into lanes 0 and 1 gives (B + 1) + 1 = B.
    ([_B, 0, 0, 0, 0, 0, 0, 0, 1, 1], _B + 2),
])
def test_dot_known_answers_fix_the_eight_lane_order(a, expected):
    """solve.rs:100-110: eight lane sums over the full chunks, combined as
    ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)), plus the tail summed in order. Every
    expected value is derived by hand in the case's comment, and each case is built so that
    the plausible other orders, named there, round to a different answer."""
    a = np.asarray(a, dtype=float)
    assert oracle.dot(a, np.ones_like(a)) == expected
    assert oracle.dot(np.ones_like(a), a) == expected
    # products, not the operands, are summed: scaling one operand by 2 and the other by 1/2
    # is exact and leaves every product, hence the dot, unchanged
    assert oracle.dot(2.0 * a, np.full_like(a, 0.5)) == expected


def test_graph_lists_each_neighbour_once():
    """iterative.rs:285-293."""
    g = oracle.Graph.from_elements(4, [[0, 1, 2], [1, 2, 3]])
    assert sorted(g.neighbours(1).tolist()) == [0, 2, 3]
    assert len(g.neighbours(0)) == 2


def _rcm_python(adj):
    """solve.rs:200-270 restated independently in Python with Rust's iterator semantics
    (min_by_key keeps the first minimum, sort_by_key is stable)."""
    n = len(adj)
    deg = [len(a) for a in adj]
    visited = [False] * n
    order = []

    def levels_from(start):
        seen = visited[:]
        seen[start] = True
        levels = [[start]]
        while True:
            nxt = []
            for v in levels[-1]:
                for w in adj[v]:
                    if not seen[w]:
                        seen[w] = True
                        nxt.append(w)
            if not nxt:
                return levels
            levels.append(nxt)

    while True:
        cand = [v for v in range(n) if not visited[v]]
        if not cand:
            break
        seed = min(cand, key=lambda v: deg[v])  # first minimum
        start, levels = seed, levels_from(seed)
        while True:
            far = min(levels[-1], key=lambda v: deg[v])
            fl = levels_from(far)
            if len(fl) <= len(levels):
                break
            start, levels = far, fl
        head = len(order)
        visited[start] = True
        order.append(start)
        while head < len(order):
            v = order[head]
            head += 1
            nxt = sorted([w for w in adj[v] if not visited[w]], key=lambda w: deg[w])
            for w in nxt:
                if not visited[w]:
                    visited[w] = True
                    order.append(w)
    return order[::-1]


def test_rcm_matches_an_independent_restatement_on_random_graphs():
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(1, 40))
        ne = int(rng.integers(0, 30))
        elements = [list(rng.choice(n, size=min(n, int(rng.integers(1, 5))), replace=False))
                    for _ in range(ne)]
        g = oracle.Graph.from_elements(n, elements)
        off, tgt = g.arrays()
        adj = [tgt[off[v]:off[v + 1]].tolist() for v in range(n)]
        # Graph::from_elements order: incident elements in order, first-seen neighbours
        expect_adj = []
        for v in range(n):
            seen, row = {v}, []
            for e in elements:
                if v in e:
                    for w in e:
                        if w not in seen:
                            seen.add(w)
                            row.append(int(w))
            expect_adj.append(row)
        assert adj == expect_adj
        assert g.rcm().tolist() == _rcm_python(adj), f"trial {trial}"


def test_rcm_is_a_permutation_that_shrinks_the_profile_of_a_shuffled_grid():
    """solve.rs:387-416."""
    side = 6
    shuffle = [3, 0, 5, 1, 4, 2]
    idx = lambda i, j: shuffle[i] * side + shuffle[j]  # noqa: E731
    adj = [[] for _ in range(side * side)]
    for i in range(side):
        for j in range(side):
            if i + 1 < side:
                adj[idx(i, j)].append(idx(i + 1, j))
                adj[idx(i + 1, j)].append(idx(i, j))
            if j + 1 < side:
                adj[idx(i, j)].append(idx(i, j + 1))
                adj[idx(i, j + 1)].append(idx(i, j))
    order = oracle.Graph.from_lists(adj).rcm()
    assert sorted(order.tolist()) == list(range(side * side))

    def profile(order):
        rank = np.empty(len(order), dtype=int)
        rank[order] = np.arange(len(order))
        return sum(rank[v] - min([rank[w] for w in adj[v]] + [rank[v]]) + 1
                   for v in range(len(adj)))

    before, after = profile(np.arange(side * side)), profile(order)
    assert after <= side * side * (side + 1) and after < before


def test_rcm_handles_disconnected_graphs_and_isolated_vertices():
    """solve.rs:418-425."""
    order = oracle.Graph.from_lists([[2], [3], [0], [1], []]).rcm()
    assert sorted(order.tolist()) == [0, 1, 2, 3, 4]


# ------------------------------------------------------------------ IC(0), PCG, skyline
def _grid_system(n):
    """iterative.rs:257-283: (4 + degree, -1) Laplacian of an n x n grid, RCM-numbered."""
    idx = lambda i, j: i + n * j  # noqa: E731
    elements = [[idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)]
                for j in range(n - 1) for i in range(n - 1)]
    g = oracle.Graph.from_elements(n * n, elements)
    order = g.rcm()
    rank = np.empty(n * n, dtype=int)
    rank[order] = np.arange(n * n)
    a = oracle.Csr.zeros(g, order)
    for v in range(n * n):
        nb = g.neighbours(v)
        a.add(rank[v], rank[v], 1.0 + len(nb))
        for w in nb[nb < v]:
            a.add(rank[v], rank[w], -1.0)
    x = np.sin(0.37 * np.arange(n * n))
    return a, x, a.multiply(x)


def test_pcg_recovers_the_chosen_solution():
    """iterative.rs:302-309."""
    a, x, b = _grid_system(30)
    st, got, it, res, _ = a.pcg(b)
    assert st == oracle.OK and np.abs(got - x).max() < 1e-8


def test_pcg_solves_a_zero_right_hand_side_exactly():
    """iterative.rs:311-324."""
    a, _, b = _grid_system(10)
    st, x, it, _, _ = a.pcg(np.zeros(100))
    assert st == oracle.OK and it == 0 and np.all(x == 0)
    st, x, _, _, _ = a.pcg(b)
    assert np.abs(a.multiply(x) - b).max() < 1e-8


def test_incomplete_cholesky_of_a_tridiagonal_matrix_is_exact():
    """iterative.rs:326-346."""
    g = oracle.Graph.from_elements(4, [[0, 1], [1, 2], [2, 3]])
    a = oracle.Csr.zeros(g, np.arange(4))
    for i in range(4):
        a.add(i, i, 2.0)
        if i > 0:
            a.add(i, i - 1, -1.0)
    l = a.incomplete_cholesky(0.0)
    z = a.precondition(l, [1.0, 0.0, 0.0, 1.0])
    assert np.abs(a.multiply(z) - [1, 0, 0, 1]).max() < 1e-13


def test_incomplete_cholesky_without_dropped_fill_is_the_cholesky_factor():
    """With a full lower pattern IC(0) has nothing to drop: it is numpy's Cholesky."""
    rng = np.random.default_rng(5)
    n = 9
    b = rng.normal(size=(n, n))
    dense = b @ b.T + n * np.eye(n)
    elements = [list(range(n))]
    g = oracle.Graph.from_elements(n, elements)
    a = oracle.Csr.zeros(g, np.arange(n))
    for i in range(n):
        for j in range(i + 1):
            a.add(i, j, dense[i, j])
    l = a.incomplete_cholesky(0.0)
    off, cols, _ = a.arrays()
    lmat = np.zeros((n, n))
    for i in range(n):
        lmat[i, cols[off[i]:off[i + 1]]] = l[off[i]:off[i + 1]]
    assert np.abs(lmat - np.linalg.cholesky(dense)).max() < 1e-12


def test_singular_chain_is_reported():
    """iterative.rs:348-362."""
    g = oracle.Graph.from_elements(3, [[0, 1], [1, 2]])
    a = oracle.Csr.zeros(g, np.arange(3))
    for i, j, v in [(0, 0, 1.0), (1, 1, 2.0), (2, 2, 1.0), (1, 0, -1.0), (2, 1, -1.0)]:
        a.add(i, j, v)
    st = a.pcg(np.ones(3))[0]
    assert st != oracle.OK


KERSHAW = np.array([[3.0, -2, 0, 2], [-2, 3, -2, 0], [0, -2, 3, -2], [2, 0, -2, 3]])


def kershaw_csr(module):
    """Kershaw's SPD matrix on a 4-cycle: unshifted IC(0) breaks down on it."""
    g = module.Graph.from_elements(4, [[0, 1], [1, 2], [2, 3], [3, 0]])
    a = module.Csr.zeros(g, np.arange(4))
    for i in range(4):
        for j in range(i + 1):
            if KERSHAW[i, j] != 0:
                a.add(i, j, KERSHAW[i, j])
    return a


def test_shift_ladder_recovers_a_failed_ic0():
    """iterative.rs:12-14, 201-205: the first shift of the ladder that factors is used."""
    assert np.linalg.eigvalsh(KERSHAW).min() > 0
    a = kershaw_csr(oracle)
    st, x, it, res, shift = a.pcg(np.array([1.0, 2.0, 3.0, 4.0]))
    assert shift > 0
    for k in range(shift):
        assert a.incomplete_cholesky(oracle.SHIFTS[k]) is None
    assert a.incomplete_cholesky(oracle.SHIFTS[shift]) is not None
    assert st == oracle.OK and np.abs(KERSHAW @ x - [1, 2, 3, 4]).max() < 1e-9


def test_skyline_solves_the_known_spd_system():
    """solve.rs:287-305."""
    a = np.array([[4, 1, 0, 0.5], [1, 5, 2, 0], [0, 2, 6, 1], [0.5, 0, 1, 3.0]])
    x = np.array([1, -2, 3, 0.5])
    s = oracle.Skyline([0, 0, 0, 0])
    for i in range(4):
        for j in range(i + 1):
            s.add(i, j, a[i, j])
    assert s.factorize() == oracle.OK
    assert np.abs(s.solve(a @ x) - x).max() < 1e-13


def test_skyline_variable_profile():
    """solve.rs:307-332."""
    n = 5
    full, sky = oracle.Skyline([0] * n), oracle.Skyline([0, 0, 1, 2, 1])
    for m in (full, sky):
        for i in range(n):
            m.add(i, i, 2.0)
            if i > 0:
                m.add(i, i - 1, -1.0)
        m.add(1, 4, 0.5)
    assert sky.stored() == 11
    b = np.array([1.0, 0.0, 2.0, -1.0, 3.0])
    full.factorize()
    sky.factorize()
    assert np.abs(full.solve(b) - sky.solve(b)).max() < 1e-13


def test_skyline_singular():
    """solve.rs:361-368."""
    s = oracle.Skyline([0, 0])
    for i, j, v in [(0, 0, 1.0), (1, 0, 1.0), (1, 1, 1.0)]:
        s.add(i, j, v)
    assert s.factorize() == oracle.SINGULAR


# ------------------------------------------------------------------ the static model
def _box_problem(n, size, quadratic, fixed_fn, imposed_fn=None):
    nodes, tets = M.box_grid(n, size)
    if quadratic:
        nodes, tets = M.to_quadratic_tet(nodes, tets)
    ne = len(tets)
    fixed = np.zeros(3 * len(nodes), dtype=bool)
    imposed = np.zeros(3 * len(nodes))
    for v, p in enumerate(nodes):
        for c in range(3):
            val = fixed_fn(p, c)
            if val is not None:
                fixed[3 * v + c] = True
                imposed[3 * v + c] = val
    return M.Problem("box", nodes, tets, np.full(ne, E), np.full(ne, NU), fixed, imposed,
                     np.zeros(3 * len(nodes)))


EPS = np.array([1e-3, -2e-3, 0.5e-3, 3e-3, -1e-3, 2e-3])


def _patch_exact(p):
    e = EPS
    return np.array([e[0] * p[0] + e[5] / 2 * p[1] + e[4] / 2 * p[2],
                     e[5] / 2 * p[0] + e[1] * p[1] + e[3] / 2 * p[2],
                     e[4] / 2 * p[0] + e[3] / 2 * p[1] + e[2] * p[2]])


def patch_problem(quadratic):
    """tests/solid.rs:52-116: linear field imposed on the boundary of a 2 x 1 x 1.5 box."""
    size = np.array([2.0, 1.0, 1.5])

    def fixed(p, c):
        on = np.any(np.abs(p) < 1e-12) or np.any(np.abs(p - size) < 1e-12)
        return _patch_exact(p)[c] if on else None

    return _box_problem((3, 2, 2), size, quadratic, fixed)


@pytest.mark.parametrize("quadratic", [False, True])
@pytest.mark.parametrize("iterative", [False, True])
def test_uniform_strain_patch_is_exact(quadratic, iterative):
    """tests/solid.rs:52-116."""
    p = patch_problem(quadratic)
    r = oracle.model_solve(p, iterative=iterative)
    assert r["status"] == oracle.OK
    exact = np.array([_patch_exact(x) for x in p.nodes]).ravel()
    assert np.abs(r["displacements"] - exact).max() < 1e-12
    sig = oracle.solid_d(E, NU) @ EPS
    assert np.abs(r["stresses"][..., 3:] - sig).max() <= 1e-9 * np.abs(sig).max()


def uniaxial_problem(quadratic, length=4.0, delta=0.01):
    def fixed(p, c):
        if abs(p[0] - length) < 1e-12 and c == 0:
            return delta
        return 0.0 if abs(p[c]) < 1e-12 else None

    return _box_problem((4, 1, 1), (length, 1.0, 1.0), quadratic, fixed)


@pytest.mark.parametrize("quadratic", [False, True])
def test_uniaxial_traction_gives_e_delta_over_l_and_poisson_contraction(quadratic):
    """tests/solid.rs:119-175."""
    length, delta = 4.0, 0.01
    p = uniaxial_problem(quadratic, length, delta)
    r = oracle.model_solve(p, iterative=False)
    strain = delta / length
    s = r["stresses"]
    assert np.abs(s[..., 3] - E * strain).max() <= 1e-9 * E * strain
    assert np.abs(s[..., 4]).max() < 1e-9 * E * strain
    u = r["displacements"].reshape(-1, 3)
    assert np.abs(u[:, 1] + NU * strain * p.nodes[:, 1]).max() < 1e-12
    assert np.abs(u[:, 2] + NU * strain * p.nodes[:, 2]).max() < 1e-12
    tip = np.flatnonzero(np.abs(p.nodes[:, 0] - length) < 1e-12)
    pull = r["reactions"][3 * tip].sum()
    assert abs(pull - E * strain) <= 1e-9 * E * strain


def test_tet10_cantilever_matches_timoshenko_beam_theory():
    """tests/solid.rs:181-248: Tet10 cantilever 10 x 1 x 1 under a unit end load, within 2%."""
    length, b, h, load = 10.0, 1.0, 1.0, 1.0
    p = M.cantilever_tet10((20, 2, 2), (length, b, h), E, NU, (0.0, 0.0, -load / (b * h)))
    r = oracle.model_solve(p, iterative=True)
    nodes = p.nodes
    centre = np.flatnonzero((np.abs(nodes[:, 0] - length) < 1e-12)
                            & (np.abs(nodes[:, 1] - b / 2) < 1e-12)
                            & (np.abs(nodes[:, 2] - h / 2) < 1e-12))[0]
    inertia, area, g = b * h ** 3 / 12, b * h, E / (2 * (1 + NU))
    theory = load * length ** 3 / (3 * E * inertia) + load * length / (5 / 6 * g * area)
    tip = -r["displacements"][3 * centre + 2]
    assert abs(tip / theory - 1) < 0.02


def test_cube_under_hydrostatic_pressure_shrinks_uniformly():
    """tests/pressure.rs:195-226: pressure p on the three faces x=1, y=1, z=1 of a unit cube on
    rollers at x=0, y=0, z=0 gives u = -p (1 - 2 nu) / E x."""
    for quadratic in (False, True):
        def fixed(p, c):
            return 0.0 if abs(p[c]) < 1e-12 else None

        prob = _box_problem((2, 2, 2), (1.0, 1.0, 1.0), quadratic, fixed)
        pres = []
        for axis in range(3):
            e, f = M.faces_on(prob.nodes, prob.conn,
                              lambda x, a=axis: np.abs(x[:, a] - 1.0) < 1e-12)
            pres += list(zip(e.tolist(), f.tolist(), [5.0] * len(e)))
        pres.sort()
        prob.pressures = pres
        r = oracle.model_solve(prob, iterative=False)
        strain = -5.0 * (1 - 2 * NU) / E
        assert np.abs(r["displacements"] - strain * prob.nodes.ravel()).max() < 1e-12


def test_iterative_static_solution_matches_the_direct_one():
    """tests/iterative.rs:69-83 on the oracle: PCG vs skyline on the Tet10 cantilever."""
    p = M.reference_cantilever()
    it = oracle.model_solve(p, iterative=True)
    di = oracle.model_solve(p, iterative=False)
    assert relative_gap(it["displacements"], di["displacements"]) < 1e-8
    assert relative_gap(it["reactions"], di["reactions"]) < 1e-8
    assert relative_gap(it["stresses"][..., 3], di["stresses"][..., 3]) < 1e-7


def test_unconstrained_structure_is_reported():
    """tests/iterative.rs:200-224."""
    p = M.reference_cantilever((4, 1, 1))
    p.fixed[:] = False
    p.f[:] = 0.0
    p.f[2] = 1.0
    r = oracle.model_solve(p, iterative=True)
    assert r["status"] in (oracle.SINGULAR, oracle.NOT_CONVERGED)


def test_workload_sizes_match_the_kuhn_enumeration():
    """Element, node, DOF and stored-entry counts of the benchmark meshes."""
    p1 = M.cantilever_tet4()
    assert (len(p1.conn), len(p1.nodes), p1.ndof, p1.nfree) == (1920, 525, 1575, 1500)
    s = oracle.System(p1.nodes, p1.conn, p1.young, p1.nu, p1.fixed)
    assert len(s.csr()[1]) == 26931
    pr = M.reference_cantilever()
    s = oracle.System(pr.nodes, pr.conn, pr.young, pr.nu, pr.fixed)
    assert len(s.csr()[1]) == 48756
    nodes, tets = M.box_grid((40, 8, 8), (40.0, 8.0, 8.0))
    nodes, tets = M.to_quadratic_tet(nodes, tets)
    assert (len(tets), len(nodes)) == (15360, 23409)


def test_jittered_mesh_is_valid_and_seeded():
    a = M.jittered_cube(6)
    b = M.jittered_cube(6)
    assert np.array_equal(a.nodes, b.nodes) and np.array_equal(a.young, b.young)
    assert M.check_tetrahedra(a.nodes, a.conn) == -1
    assert abs(np.median(np.log(M.jittered_cube(20).young / 200000.0))) < 0.02


def test_box_grid_matches_a_literal_loop_restatement():
    """tests/common/mod.rs:74-114 written as its loops, on a small box."""
    n, size = (3, 2, 2), (3.0, 2.0, 1.0)
    nodes, tets = M.box_grid(n, size)
    idx = lambda p: p[0] + (n[0] + 1) * (p[1] + (n[1] + 1) * p[2])  # noqa: E731
    expect = []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                for axes in itertools.permutations(range(3)):
                    p = [i, j, k]
                    tet = [idx(p)]
                    for a in axes:
                        p[a] += 1
                        tet.append(idx(p))
                    c = nodes[tet]
                    if np.linalg.det(c[1:] - c[0]) < 0:
                        tet[1], tet[2] = tet[2], tet[1]
                    expect.append(tet)
    assert np.array_equal(tets, np.array(expect))
    q_nodes, q = M.to_quadratic_tet(nodes, tets)
    mids, order = {}, len(nodes)
    for t in tets:
        for a, b in M.TET10_EDGES:
            key = (min(t[a], t[b]), max(t[a], t[b]))
            if key not in mids:
                mids[key] = order
                order += 1
    for e, t in enumerate(tets):
        assert [mids[(min(t[a], t[b]), max(t[a], t[b]))] for a, b in M.TET10_EDGES] == q[e, 4:].tolist()


# ------------------------------------------------------------------ adjoint compliance gradient
def test_compliance_gradient_matches_central_differences():
    """The CPU adjoint dC/dE_e = -u_e^T (K_e/E_e) u_e against central differences of the
    compliance C = f.u computed with the skyline direct solver (solve.rs:10-139), on a small
    SIMP box: the derivative of each probed element's modulus agrees to 1e-6."""
    p = M.simp_box((6, 3, 3), (6.0, 3.0, 3.0))
    base = oracle.model_solve(p, iterative=False)
    assert base["status"] == oracle.OK
    u = base["displacements"]
    g = oracle.compliance_gradient(p.nodes, p.conn, p.young, p.nu, u)
    assert np.all(g < 0)  # stiffening any element lowers the compliance
    rng = np.random.default_rng(6)
    for e in rng.choice(len(p.conn), 8, replace=False):
        h = 1e-3 * p.young[e]
        cs = []
        for sgn in (1.0, -1.0):
            q = M.Problem(p.name, p.nodes, p.conn, p.young.copy(), p.nu, p.fixed, p.imposed,
                          p.f, p.pressures, p.meta)
            q.young[e] += sgn * h
            r = oracle.model_solve(q, iterative=False)
            cs.append(float(p.f @ r["displacements"]))
        fd = (cs[0] - cs[1]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * abs(g[e]), (e, fd, g[e])
    # compliance is the sum of the element energies: C = -sum_e E_e dC/dE_e
    assert abs(-(p.young * g).sum() - p.f @ u) <= 1e-10 * abs(p.f @ u)


# ------------------------------------------------------------------ nodal stresses
TET10_REFERENCE_NODES = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0.5, 0, 0),
                                  (0.5, 0.5, 0), (0, 0.5, 0), (0, 0, 0.5), (0.5, 0, 0.5),
                                  (0, 0.5, 0.5)], dtype=float)


def test_tet10_extrapolation_reproduces_a_linear_function_at_the_ten_nodes():
    """results.rs:263-285, through nodal_stresses on the reference Tet10 (one element, so no
    averaging): f = 2 + 3 xi - 5 eta + 7 zeta known at the TETRA_4 points (quadrature.rs:108-138,
    point g near vertex g) comes out at the ten nodes, in each of the six components."""
    a, b = 0.585410196624968500, 0.138196601125010500
    points = [(b, b, b), (a, b, b), (b, a, b), (b, b, a)]
    f = lambda x, y, z: 2.0 + 3.0 * x - 5.0 * y + 7.0 * z
    scale = np.array([1.0, -2.0, 3.0, 0.5, 4.0, -7.0])
    gp = np.zeros((1, 4, 9))
    for g, p in enumerate(points):
        gp[0, g, :3] = p
        gp[0, g, 3:] = f(*p) * scale
    out = oracle.nodal_stresses(TET10_REFERENCE_NODES, np.arange(10)[None, :], gp)
    assert np.array_equal(out[:, :3], TET10_REFERENCE_NODES)
    for k, node in enumerate(TET10_REFERENCE_NODES):
        assert np.abs(out[k, 3:] - f(*node) * scale).max() < 1e-12, k


def test_nodal_stresses_average_over_the_elements_at_a_node():
    """results.rs:200-220: each node takes the mean over the elements that use it (here two Tet4
    sharing a face, each with its own constant stress); a node no element uses keeps zeros."""
    nodes = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (5, 5, 5)], float)
    conn = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    s0 = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    s1 = np.array([-3.0, 0.25, 7.0, 1.5, -2.0, 10.0])
    gp = np.zeros((2, 1, 9))
    gp[0, 0, 3:], gp[1, 0, 3:] = s0, s1
    out = oracle.nodal_stresses(nodes, conn, gp)
    assert np.array_equal(out[:, :3], nodes)
    assert np.array_equal(out[0, 3:], s0) and np.array_equal(out[4, 3:], s1)
    for n in (1, 2, 3):
        assert np.array_equal(out[n, 3:], (s0 + s1) / 2)
    assert np.array_equal(out[5, 3:], np.zeros(6))


@pytest.mark.parametrize("nen", [4, 10])
def test_nodal_stresses_of_a_linear_field_are_d_eps(nen):
    """A linear displacement field on a box mesh: every Gauss point carries D eps
    (element.rs:840-860), so every node, after extrapolation and averaging, carries it too."""
    nodes, tets = M.box_grid((3, 2, 2), (3.0, 2.0, 2.0))
    if nen == 10:
        nodes, tets = M.to_quadratic_tet(nodes, tets)
    eps = np.array([1e-3, -2e-3, 0.5e-3, 3e-3, -1e-3, 2e-3])
    grad = np.array([[eps[0], eps[5] / 2, eps[4] / 2],
                     [eps[5] / 2, eps[1], eps[3] / 2],
                     [eps[4] / 2, eps[3] / 2, eps[2]]])
    u = (nodes @ grad.T).ravel()
    gp = oracle.stresses(nodes, tets, E, NU, u)
    out = oracle.nodal_stresses(nodes, tets, gp)
    sig = oracle.solid_d(E, NU) @ eps
    assert np.abs(out[:, 3:] - sig).max() < 1e-9 * np.abs(sig).max()


# ------------------------------------------------------------------ edge loads
TET_EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]


def _edges_where(nodes, conn, on):
    """tests/edge_load.rs:22-39: the (element, edge) pairs all of whose nodes satisfy `on`, once
    per pair of vertices, in element order."""
    seen, found = set(), []
    nen = conn.shape[1]
    for i, e in enumerate(conn):
        for k, (a, b) in enumerate(TET_EDGES):
            ids = [e[a], e[b]] + ([e[4 + k]] if nen == 10 else [])
            key = (min(e[a], e[b]), max(e[a], e[b]))
            if all(on(nodes[v]) for v in ids) and key not in seen:
                seen.add(key)
                found.append((i, k))
    return found


@pytest.mark.parametrize("nen", [4, 10])
def test_the_resultant_on_a_tetrahedron_edge_splits_as_the_shape_functions(nen):
    """tests/edge_load.rs:115-138 (Tet4: q L / 2 on each end) and :330-360 (Tet10: q L times 1/6,
    1/6, 2/3 on the two ends and the mid-node), every edge of the first element of a 1x1x1 box
    of 1 x 2 x 2."""
    nodes, tets = M.box_grid((1, 1, 1), (1.0, 2.0, 2.0))
    if nen == 10:
        nodes, tets = M.to_quadratic_tet(nodes, tets)
    q = np.array([1.0, -2.0, 3.0])
    share = [0.5, 0.5] if nen == 4 else [1 / 6, 1 / 6, 2 / 3]
    for k, (a, b) in enumerate(TET_EDGES):
        f = oracle.add_edge_loads(nodes, tets, [(0, k, tuple(q))], np.zeros(3 * len(nodes)))
        ids = [tets[0][a], tets[0][b]] + ([tets[0][4 + k]] if nen == 10 else [])
        length = np.sqrt(((nodes[ids[1]] - nodes[ids[0]]) ** 2).sum())
        for v, w in zip(ids, share):
            assert np.allclose(f[3 * v:3 * v + 3], q * length * w, rtol=1e-12, atol=0), k
        assert np.allclose(f.reshape(-1, 3).sum(axis=0), q * length, rtol=1e-12, atol=0)
        others = np.setdiff1d(np.arange(len(nodes)), ids)
        assert not f.reshape(-1, 3)[others].any()


@pytest.mark.parametrize("quadratic", [False, True])
def test_the_edge_load_resultant_is_q_times_the_rim_length(quadratic):
    """tests/edge_load.rs:291-328 (small strains): q along x on the four edges of the face x = 1
    of a unit cube on rollers; the x reactions sum to -4 q."""
    def fixed(p, c):
        return 0.0 if abs(p[c]) < 1e-9 else None

    prob = _box_problem((2, 2, 2), (1.0, 1.0, 1.0), quadratic, fixed)
    end = lambda x: abs(x[0] - 1.0) < 1e-9
    rim = []
    for side in (lambda x: abs(x[1]) < 1e-9, lambda x: abs(x[1] - 1.0) < 1e-9,
                 lambda x: abs(x[2]) < 1e-9, lambda x: abs(x[2] - 1.0) < 1e-9):
        rim += _edges_where(prob.nodes, prob.conn, lambda x, s=side: end(x) and s(x))
    q = 25.0
    prob.edge_loads = [(e, k, (q, 0.0, 0.0)) for e, k in rim]
    r = oracle.model_solve(prob, iterative=False)
    assert abs(r["reactions"][0::3].sum() + 4.0 * q) < 1e-9 * (1.0 + 4.0 * q)
