"""Host-side checks of the HIP library (no device needed): it loads and exports every entry
point its header declares, and the index layouts its kernels read are in bounds, cover every
entry of L once in the reference's order, and give the oracle's preconditioner bit for bit
when applied block by block in the device's processing order."""
import ctypes
import os
import re

import numpy as np
import pytest

import maille_amd as ma
import oracle
from maille_amd import meshes as M
from test_oracle import kershaw_csr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "maille_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ml_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    names = header_functions()
    assert len(names) > 30
    lib = ctypes.CDLL(ma._LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(ma.EXPORTS)
    ma.lib()  # every signature binds


def _problem_csr(p):
    s = oracle.System(p.nodes, p.conn, p.young, p.nu, p.fixed, p.imposed)
    off, cols, vals = s.csr()
    assert s.factorize() == oracle.OK
    return s, off, cols, vals


def _chain(n, bandwidth):
    """Lower CSR of a banded SPD matrix (ascending columns, diagonal last)."""
    off, cols = [0], []
    for i in range(n):
        cols += list(range(max(0, i - bandwidth), i)) + [i]
        off.append(len(cols))
    return np.array(off), np.array(cols)


PROBLEMS = {
    "reference_cantilever": lambda: M.reference_cantilever(),
    "cantilever_tet4": lambda: M.cantilever_tet4(),
    "jittered_cube_8": lambda: M.jittered_cube(8),
    "cantilever_tet10_small": lambda: M.cantilever_tet10((12, 3, 3), (60.0, 15.0, 15.0)),
}


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_layouts_of_mesh_matrices(name):
    s, off, cols, vals = _problem_csr(PROBLEMS[name]())
    l = s.l_values()
    r = np.cos(0.11 * np.arange(s.n))
    stats, z = ma.layout_verify(off, cols, l, r)
    assert stats["fw_blocks"] > 1 and stats["bw_blocks"] > 1
    assert 1 <= stats["depth"] <= s.n and stats["max_row"] <= 512
    ref = oracle.Csr.from_arrays(off, cols, vals).precondition(l, r)
    assert np.array_equal(z, ref)


def test_layouts_of_a_single_block_without_external_entries():
    """Kershaw's 4x4 matrix fits in one block: both solves have no external entries at all
    (the device then gathers nothing for them)."""
    a = kershaw_csr(oracle)
    off, cols, vals = a.arrays()
    stats = ma.layout_verify(off, cols)
    assert stats["fw_blocks"] == 1 and stats["bw_blocks"] == 1
    assert stats["fw_max_ext"] == 0 and stats["bw_max_ext"] == 0
    l = next(f for f in (a.incomplete_cholesky(sh) for sh in oracle.SHIFTS) if f is not None)
    r = np.array([1.0, -2.0, 0.5, 3.0])
    _, z = ma.layout_verify(off, cols, l, r)
    assert np.array_equal(z, a.precondition(l, r))


@pytest.mark.parametrize("n,bandwidth", [(1, 0), (7, 0), (200, 1), (300, 5), (130, 40)])
def test_layouts_of_banded_matrices(n, bandwidth):
    """Diagonal, chain and banded patterns: blocks of at most 64 rows cut where the chain
    breaks, internal entries only to earlier rows of the block."""
    off, cols = _chain(n, bandwidth)
    rng = np.random.default_rng(n + bandwidth)
    l = rng.uniform(-0.3, 0.3, len(cols))
    l[off[1:] - 1] = rng.uniform(1.0, 2.0, n)
    r = rng.normal(size=n)
    stats, z = ma.layout_verify(off, cols, l, r)
    assert stats["max_row"] == bandwidth + 1
    if bandwidth == 0:
        assert stats["fw_max_ext"] == 0 and stats["fw_max_int"] == 0
    # the same solve, row by row (iterative.rs:175-186)
    L = np.zeros((n, n))
    for i in range(n):
        L[i, cols[off[i]:off[i + 1]]] = l[off[i]:off[i + 1]]
    y = np.zeros(n)
    for i in range(n):
        acc = 0.0
        for p in range(off[i], off[i + 1] - 1):
            acc = acc + l[p] * y[cols[p]]
        y[i] = (r[i] - acc) / l[off[i + 1] - 1]
    zz = np.zeros(n)
    for i in reversed(range(n)):
        acc = y[i]
        for k in reversed(range(i + 1, n)):
            if L[k, i] != 0.0:
                acc = acc - L[k, i] * zz[k]
        zz[i] = acc / L[i, i]
    assert np.array_equal(z, zz)


def test_row_longer_than_the_block_capacity_is_rejected():
    """A row of L with more off-diagonal entries than a block can stage is refused on the
    host with a clear error instead of reaching the device."""
    n = 700
    off, cols = [0], []
    for i in range(n - 1):
        cols.append(i)
        off.append(len(cols))
    cols += list(range(n))  # last row: dense
    off.append(len(cols))
    with pytest.raises(ValueError, match="longer than the triangular-solve block capacity"):
        ma.layout_verify(np.array(off), np.array(cols))


def test_malformed_csr_is_rejected():
    off, cols = _chain(10, 2)
    bad = cols.copy()
    bad[5] = 9  # a column after the diagonal
    with pytest.raises(ValueError):
        ma.layout_verify(off, bad)
