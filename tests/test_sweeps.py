"""The Jacobi-sweep restatement (tests/sweeps.py) pinned by what the mathematics fixes, and
the library's sweep layouts checked on the host."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle
from maille_amd import meshes
from sweeps import jacobi_precondition


def _factored(p):
    s = oracle.System(p.nodes, p.conn, p.young, p.nu, p.fixed)
    assert s.factorize() == oracle.OK
    off, cols, vals = s.csr()
    return off, cols, s.l_values(), s.n


def _depth(off, cols):
    n = len(off) - 1
    level = np.zeros(n, dtype=np.int64)
    for i in range(n):
        js = cols[off[i]:off[i + 1] - 1]
        level[i] = level[js].max() + 1 if len(js) else 0
    return int(level.max()) + 1


def test_enough_sweeps_are_the_exact_triangular_solves():
    """Jacobi on a triangular system is exact after as many sweeps as its DAG has levels
    (row i is final after level(i) + 1 sweeps): the restatement then equals the reference's
    exact precondition to rounding. A transposed triangle, a wrong row sum or a wrong
    diagonal would not converge to it."""
    p = meshes.cantilever_tet4(n=(6, 2, 2), size=(30.0, 10.0, 10.0))
    off, cols, l, n = _factored(p)
    r = np.sin(0.7 * np.arange(n)) + 0.1
    exact = oracle.Csr.from_arrays(off, cols, np.zeros(len(cols))).precondition(l, r)
    k = _depth(off, cols)
    z = jacobi_precondition(off, cols, l, r, k)
    assert np.linalg.norm(z - exact) <= 1e-12 * np.linalg.norm(exact)
    # one sweep is still far from it; the error falls with every sweep
    errs = [np.linalg.norm(jacobi_precondition(off, cols, l, r, m) - exact) for m in (1, 2, 4)]
    assert errs[0] > 1e-3 * np.linalg.norm(exact)
    assert errs[0] > errs[1] > errs[2]


def test_sweeps_are_the_polynomial_preconditioner():
    """k forward sweeps from y = D^-1 r are y = sum_{m<=k} (-D^-1 N)^m D^-1 r =: P r, the
    backward ones z = P^T y, so M^-1 = P^T P: checked against scipy's sparse products, and
    M^-1 is symmetric."""
    p = meshes.jittered_cube(4)
    off, cols, l, n = _factored(p)
    rows = np.repeat(np.arange(n), np.diff(off))
    L = sp.csr_matrix((l, (rows, cols)), shape=(n, n))
    d = L.diagonal()
    N = sp.tril(L, -1).tocsr()
    Dinv = sp.diags(1.0 / d)
    rng = np.random.default_rng(3)
    for k in (1, 2, 5):
        def P(v):
            out = Dinv @ v
            term = out
            for _ in range(k):
                term = -(Dinv @ (N @ term))
                out = out + term
            return out

        def PT(v):
            out = Dinv @ v
            term = out
            for _ in range(k):
                term = -(Dinv @ (N.T @ term))
                out = out + term
            return out
        r = rng.standard_normal(n)
        z = jacobi_precondition(off, cols, l, r, k)
        ref = PT(P(r))
        assert np.linalg.norm(z - ref) <= 1e-10 * np.linalg.norm(ref)
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        ma = jacobi_precondition(off, cols, l, a, k)
        mb = jacobi_precondition(off, cols, l, b, k)
        assert abs(b @ ma - a @ mb) <= 1e-10 * abs(a @ mb)
        assert a @ ma > 0.0


def test_sweeps_of_a_diagonal_factor_are_the_diagonal_solve():
    """With no off-diagonal entries every sweep is the exact D^-2 r."""
    n = 5
    off = np.arange(n + 1)
    cols = np.arange(n)
    l = np.array([2.0, 4.0, 0.5, 8.0, 1.0])
    r = np.array([1.0, -3.0, 2.0, 5.0, 7.0])
    for k in (1, 3):
        assert np.array_equal(jacobi_precondition(off, cols, l, r, k), r * (1 / l) * (1 / l))


def test_library_sweep_layouts_pass_their_host_checks():
    """ml_layout_verify builds and checks the sweep layouts with the other device layouts."""
    import maille_amd as ma
    p = meshes.jittered_cube(5)
    off, cols, l, n = _factored(p)
    stats = ma.layout_verify(off, cols)
    assert stats["depth"] > 1
