"""Float64 restatement of the Jacobi-sweep preconditioner of ML_MIXED_JACOBI (no counterpart in
the reference, whose Csr::precondition, iterative.rs:173-188, solves both triangular systems
exactly). Test infrastructure: the device result is compared with it bit for bit.

With rd = 1 / l_ii (correctly rounded) and N the strict lower part of L (lower CSR,
ascending columns, diagonal last):
    y = r * rd, then k times  y_i <- (r_i - sum_{j<i} l_ij y_j) * rd_i
    z = y * rd, then k times  z_i <- (y_i - sum_{j>i} l_ji z_j) * rd_i
each row sum taken sequentially in ascending j from 0, every operation rounded to float64.
The k - 1 intermediate iterates of each system are the values the fused kernel exchanges
between workgroups with a launch tag in their low 8 significand bits: they are rounded toward
zero to 44 fraction bits (those bits cleared); the last iterate of each system is not.
With a weight w != 1 every sweep is damped: y_i <- c1 * y_i + w * t_i, t_i the undamped
result and c1 = 1 - w, each product and the sum rounded to float64 in that order.
"""
import numpy as np


def _padded_rows(n, rows, cols, vals):
    """(n, width) arrays of each row's entries in ascending column, and a validity mask."""
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    counts = np.bincount(rows, minlength=n)
    width = int(counts.max(initial=0))
    start = np.zeros(n + 1, dtype=np.int64)
    start[1:] = np.cumsum(counts)
    pos = np.arange(len(rows)) - start[rows]
    c = np.zeros((n, max(width, 1)), dtype=np.int64)
    v = np.zeros((n, max(width, 1)))
    m = np.zeros((n, max(width, 1)), dtype=bool)
    c[rows, pos] = cols
    v[rows, pos] = vals
    m[rows, pos] = True
    return c, v, m


def _sweep(c, v, m, rhs, x, rd, w=1.0):
    s = np.zeros(len(rhs))
    for q in range(c.shape[1]):
        s = np.where(m[:, q], s + v[:, q] * x[c[:, q]], s)
    t = (rhs - s) * rd
    return t if w == 1.0 else (1.0 - w) * x + w * t


def level_value(v):
    """v with the low 8 bits of its float64 significand cleared."""
    return (np.asarray(v, dtype=np.float64).view(np.uint64) & np.uint64(0xFFFFFFFFFFFFFF00)).view(np.float64)


def jacobi_precondition(off, cols, l, r, sweeps, w=1.0):
    off = np.asarray(off, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    l = np.asarray(l, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    n = len(off) - 1
    diag = off[1:] - 1
    rd = 1.0 / l[diag]
    rows = np.repeat(np.arange(n), np.diff(off))
    strict = np.ones(len(cols), dtype=bool)
    strict[diag] = False
    lr, lc, lv = rows[strict], cols[strict], l[strict]
    lo = _padded_rows(n, lr, lc, lv)
    up = _padded_rows(n, lc, lr, lv)  # rows of L^T: column j of L, ascending row
    y = r * rd
    for m in range(1, sweeps + 1):
        y = _sweep(*lo, r, y, rd, w)
        if m < sweeps:
            y = level_value(y)
    z = y * rd
    for m in range(1, sweeps + 1):
        z = _sweep(*up, y, z, rd, w)
        if m < sweeps:
            z = level_value(z)
    return z
