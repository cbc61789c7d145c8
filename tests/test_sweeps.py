"""The Jacobi-sweep restatement (tests/sweeps.py) pinned by what the mathematics fixes, and
the library's sweep layouts checked on the host."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle
from maille_amd import meshes
from sweeps import jacobi_precondition, level_value


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


def test_damped_sweeps_are_a_symmetric_positive_definite_preconditioner():
    """With a weight w < 1 each sweep is y <- (1 - w) y + w D^-1 (r - N y) (and the transposed
    iteration on L^T): checked against scipy's sparse products; M^-1 stays symmetric and
    positive definite; and since the iteration matrix (1 - w) I - w D^-1 N has spectral radius
    1 - w, many sweeps converge to the exact triangular solves, as with w = 1."""
    p = meshes.jittered_cube(4)
    off, cols, l, n = _factored(p)
    rows = np.repeat(np.arange(n), np.diff(off))
    L = sp.csr_matrix((l, (rows, cols)), shape=(n, n))
    d = L.diagonal()
    N = sp.tril(L, -1).tocsr()
    rng = np.random.default_rng(7)
    exact = oracle.Csr.from_arrays(off, cols, np.zeros(len(cols)))
    for w in (0.8, 0.6):
        for k in (1, 3):
            r = rng.standard_normal(n)
            y = r / d
            for _ in range(k):
                y = (1 - w) * y + w * (r - N @ y) / d
            z = y / d
            for _ in range(k):
                z = (1 - w) * z + w * (y - N.T @ z) / d
            got = jacobi_precondition(off, cols, l, r, k, w)
            assert np.linalg.norm(got - z) <= 1e-12 * np.linalg.norm(z)
            a, b = rng.standard_normal(n), rng.standard_normal(n)
            ma = jacobi_precondition(off, cols, l, a, k, w)
            mb = jacobi_precondition(off, cols, l, b, k, w)
            assert abs(b @ ma - a @ mb) <= 1e-10 * abs(a @ mb)
            assert a @ ma > 0.0
        r = np.sin(0.3 * np.arange(n))
        ref = exact.precondition(l, r)
        far = jacobi_precondition(off, cols, l, r, 200, w)
        assert np.linalg.norm(far - ref) <= 1e-10 * np.linalg.norm(ref)
    # w = 1 is the undamped sweep exactly
    r = rng.standard_normal(n)
    assert np.array_equal(jacobi_precondition(off, cols, l, r, 3, 1.0),
                          jacobi_precondition(off, cols, l, r, 3))


def test_exchanged_iterates_are_rounded_toward_zero_to_44_fraction_bits():
    """The rounding of the iterates the fused sweeps exchange clears the low 8 bits of the
    significand: toward zero in magnitude for either sign, values on the 2^-44 grid unchanged,
    and at most 2^-44 relative. Known answers written from the definition (ulp 2^-52 at 1)."""
    u = 2.0 ** -52
    assert level_value(1.0 + 255 * u) == 1.0
    assert level_value(1.0 + 257 * u) == 1.0 + 256 * u
    assert level_value(-(1.0 + 511 * u)) == -(1.0 + 256 * u)
    assert level_value(3.0 * 2.0 ** 40) == 3.0 * 2.0 ** 40
    assert np.signbit(level_value(-0.0)) and level_value(-0.0) == 0.0
    rng = np.random.default_rng(5)
    v = rng.standard_normal(1000) * 10.0 ** rng.integers(-30, 30, 1000)
    w = level_value(v)
    assert np.all(np.abs(w) <= np.abs(v)) and np.all(np.sign(w) == np.sign(v))
    assert np.all(np.abs(v - w) < 2.0 ** -44 * np.abs(v))
    assert np.array_equal(level_value(w), w)


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


def _fp32_factor(p):
    off, cols, l, n = _factored(p)
    return off, cols, l.astype(np.float32).astype(np.float64), n


def test_fused_layouts_reproduce_the_restatement_on_the_host():
    """The fused kernel's layouts (blocks of 256 or 512 rows, each row's entries as LDS slots of
    the block's own rows and of its distinct external columns), applied on the host the way
    the kernel stages them, give the restatement bit for bit: every slot names the right value.
    The Tet10 mesh has rows of up to 146 strict entries and blocks reading many external
    columns."""
    import maille_amd as ma
    for p in (meshes.jittered_cube(8),
              meshes.cantilever_tet10((12, 3, 3), (60.0, 15.0, 15.0))):
        off, cols, l, n = _fp32_factor(p)
        r = np.cos(0.11 * np.arange(n))
        for rows in (256, 512):
            for k in (1, 2, 3, 5):
                stats, z = ma.layout_sweeps(off, cols, l, r, k, rows=rows)
                assert np.array_equal(z, jacobi_precondition(off, cols, l, r, k)), (p.name, rows, k)
            assert stats["blocks"] == (n + rows - 1) // rows
        assert 0 < stats["max_ext_lower"] <= 1792 and 0 < stats["max_ext_upper"] <= 1792
        widths = np.diff(off) - 1
        assert stats["width"] == max(widths.max(), np.bincount(
            cols[np.setdiff1d(np.arange(len(cols)), off[1:] - 1)], minlength=n).max())


def test_fused_layout_refuses_a_row_too_long_for_it():
    """A row of more than 252 strict entries does not fit the fused kernel's 8-bit row
    lengths: the layout is refused with the reason, and the library then keeps one launch per
    sweep (ml_layout_verify still accepts the pattern)."""
    import maille_amd as ma
    n = 254
    # an arrow: the last row couples to every other row (253 strict entries, within the
    # triangular solves' capacity of 256 distinct columns per block)
    off = np.zeros(n + 1, dtype=np.int64)
    cols = []
    for i in range(n):
        row = [i] if i < n - 1 else list(range(n))
        cols += row
        off[i + 1] = len(cols)
    cols = np.array(cols, dtype=np.int64)
    with pytest.raises(ValueError, match="too long"):
        ma.layout_sweeps(off, cols)
    assert ma.layout_verify(off, cols)["max_row"] == n


def test_every_sweep_pattern_starts_on_a_variant_that_exists():
    """The fused variant a handle starts with before autotuning, chosen on the host from the
    pattern: one row per lane with every row in registers on the elasticity cube and, for the
    one-DOF-per-node (thermal) pattern of the same mesh, whose rows hold at most 16 strict
    entries, the only variant of that width (116, compiled for a bounded register count):
    never a variant outside the table."""
    import maille_amd as ma
    p = meshes.jittered_cube(6)
    off, cols, l, n = _factored(p)
    stats = ma.layout_sweeps(off, cols)
    assert stats["default_variant"] in (124, 132, 148)
    assert stats["width"] <= stats["default_variant"] % 100
    # one DOF per node: the node graph's lower pattern in its RCM order (Csr::zeros)
    g = ma.Graph.from_elements(len(p.nodes), [list(e) for e in p.conn])
    soff, scols = g.lower_pattern()
    st = ma.layout_sweeps(soff, scols)
    assert st["width"] <= 16
    assert st["default_variant"] == 116
