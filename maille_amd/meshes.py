"""Structured tetrahedral workloads.

Restates the reference's mesh helpers (crates/maille-core/tests/common/mod.rs) vectorised in
numpy, with the same node and element numbering, and builds the solve problems the benches
and tests use (clamped faces, consistent traction loads, face pressures, per-element moduli).

Randomness comes from SplitMix64 streams so a workload is fixed by its seed on any machine.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# shape.rs:129 (faces, outward) and :135 (Tet10 mid-node edges)
TET_FACES = np.array([[0, 1, 2], [0, 3, 1], [1, 3, 2], [2, 3, 0]], dtype=np.int64)
TET10_EDGES = np.array([[0, 1], [1, 2], [2, 0], [0, 3], [1, 3], [2, 3]], dtype=np.int64)
# tests/solid.rs:190-196: faces (vertices, mid-nodes) used for consistent traction loads
TRACTION_FACES = (
    ((0, 1, 2), (4, 5, 6)),
    ((0, 1, 3), (4, 8, 7)),
    ((1, 2, 3), (5, 9, 8)),
    ((0, 2, 3), (6, 9, 7)),
)


@dataclass
class Problem:
    """A static linear-elastic problem in the reference's terms (model.rs:230-247): nodes,
    Tet4/Tet10 connectivity with a material per element, fixed DOFs with imposed values,
    nodal loads already summed into ``f``, dead face pressures (element, face, value) and dead
    line loads on edges (element, edge, (qx, qy, qz))."""

    name: str
    nodes: np.ndarray          # (N, 3) float64
    conn: np.ndarray           # (ne, 4|10) int64
    young: np.ndarray          # (ne,) float64
    nu: np.ndarray             # (ne,) float64
    fixed: np.ndarray          # (3N,) bool
    imposed: np.ndarray        # (3N,) float64
    f: np.ndarray              # (3N,) float64, nodal loads
    pressures: list = field(default_factory=list)  # [(element, face, value)]
    meta: dict = field(default_factory=dict)
    edge_loads: list = field(default_factory=list)  # [(element, edge, (qx, qy, qz))]

    @property
    def nen(self) -> int:
        return self.conn.shape[1]

    @property
    def ndof(self) -> int:
        return 3 * len(self.nodes)

    @property
    def nfree(self) -> int:
        return int(self.ndof - np.count_nonzero(self.fixed))


# ---------------------------------------------------------------- random streams
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """The first ``count`` outputs of SplitMix64 started at ``seed`` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, count: int) -> np.ndarray:
    """Standard normals by Box-Muller on consecutive uniform pairs (one normal per pair)."""
    u = uniform(seed, 2 * count).reshape(count, 2)
    return np.sqrt(-2.0 * np.log1p(-u[:, 0])) * np.cos(2.0 * np.pi * u[:, 1])


# ---------------------------------------------------------------- meshes
def box_grid(n, size):
    """tests/common/mod.rs:74-114 — box of nx*ny*nz cubes, each cut into 6 Kuhn tetrahedra,
    oriented like Tet4. Node (i, j, k) has index i + (nx+1)(j + (ny+1)k)."""
    nx, ny, nz = (int(v) for v in n)
    sx, sy, sz = (float(v) for v in size)
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1),
                             indexing="ij")
    nodes = np.stack([(sx * ii.astype(np.float64)) / nx,
                      (sy * jj.astype(np.float64)) / ny,
                      (sz * kk.astype(np.float64)) / nz], axis=-1).reshape(-1, 3)
    perms = np.array([[0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]])
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([ci.ravel(), cj.ravel(), ck.ravel()], axis=-1).astype(np.int64)
    ncell = base.shape[0]
    p = np.repeat(base[:, None, :], 6, axis=1)  # (ncell, 6, 3)
    tets = np.empty((ncell, 6, 4), dtype=np.int64)

    def idx(q):
        return q[..., 0] + (nx + 1) * (q[..., 1] + (ny + 1) * q[..., 2])

    tets[..., 0] = idx(p)
    for v in range(3):
        p = p.copy()
        p[np.arange(ncell)[:, None], np.arange(6)[None, :], perms[None, :, v]] += 1
        tets[..., v + 1] = idx(p)
    tets = tets.reshape(-1, 4)
    neg = signed_volume(nodes, tets) < 0.0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return nodes, tets


def signed_volume(nodes, tets):
    """tests/common/mod.rs:116-124 — 6V of each tetrahedron (its first four nodes)."""
    c = nodes[tets[:, :4]]
    a, b, d = c[:, 1] - c[:, 0], c[:, 2] - c[:, 0], c[:, 3] - c[:, 0]
    return (a[:, 0] * (b[:, 1] * d[:, 2] - b[:, 2] * d[:, 1])
            - a[:, 1] * (b[:, 0] * d[:, 2] - b[:, 2] * d[:, 0])
            + a[:, 2] * (b[:, 0] * d[:, 1] - b[:, 1] * d[:, 0]))


def to_quadratic_tet(nodes, tets):
    """tests/common/mod.rs:141-165 — Tet10 mesh: one node at the middle of each edge, shared
    between neighbours, appended in first-seen order (element order, then TET10_EDGES order)."""
    n = len(nodes)
    ends = tets[:, TET10_EDGES]  # (ne, 6, 2)
    lo = ends.min(axis=2).ravel()
    hi = ends.max(axis=2).ravel()
    keys = lo * n + hi
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(uniq))
    mid = (n + rank[inv]).reshape(-1, 6)
    by_rank = np.argsort(rank)
    a = lo[first[by_rank]]
    b = hi[first[by_rank]]
    mids = (nodes[a] + nodes[b]) / 2.0
    return np.concatenate([nodes, mids]), np.concatenate([tets, mid], axis=1)


def check_tetrahedra(nodes, conn):
    """mesh.rs:329-348 (check_tetrahedron): index of the first degenerate or inverted
    tetrahedron, or -1."""
    c = nodes[conn[:, :4]]
    e = c[:, 1:4] - c[:, 0:1]
    six_v = (e[:, 0, 0] * (e[:, 1, 1] * e[:, 2, 2] - e[:, 1, 2] * e[:, 2, 1])
             - e[:, 0, 1] * (e[:, 1, 0] * e[:, 2, 2] - e[:, 1, 2] * e[:, 2, 0])
             + e[:, 0, 2] * (e[:, 1, 0] * e[:, 2, 1] - e[:, 1, 1] * e[:, 2, 0]))
    longest = np.zeros(len(conn))
    for a, b in TET10_EDGES:
        longest = np.maximum(longest, np.sqrt(((c[:, a] - c[:, b]) ** 2).sum(axis=1)))
    bad = (np.abs(six_v) <= 1e-12 * longest ** 3) | (six_v < 0.0)
    idx = np.flatnonzero(bad)
    return int(idx[0]) if len(idx) else -1


def faces_on(nodes, conn, predicate):
    """(element, face) pairs, in element then face order (TET_FACES), whose three vertices
    all satisfy ``predicate(xyz) -> bool array``."""
    on = predicate(nodes)
    hit = on[conn[:, TET_FACES]].all(axis=2)  # (ne, 4)
    e, f = np.nonzero(hit)
    return e, f


def traction_loads(nodes, conn, predicate, traction, ndof):
    """tests/solid.rs:188-229 — consistent nodal loads of a uniform traction on the faces whose
    vertices satisfy ``predicate``: A/3 per vertex (Tet4), A/3 per mid-node and 0 at the
    vertices (Tet10). The face area is the norm of the edge cross product over 2."""
    quadratic = conn.shape[1] == 10
    on = predicate(nodes)
    f = np.zeros(ndof)
    t = np.asarray(traction, dtype=np.float64)
    for vertices, mids in TRACTION_FACES:
        hit = np.flatnonzero(on[conn[:, list(vertices)]].all(axis=1))
        if len(hit) == 0:
            continue
        p = nodes[conn[np.ix_(hit, list(vertices))]]
        cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        area = np.sqrt((cr ** 2).sum(axis=1)) / 2.0
        carriers = conn[np.ix_(hit, list(mids if quadratic else vertices))]
        for c in range(3):
            if t[c] != 0.0:
                np.add.at(f, 3 * carriers.ravel() + c, np.repeat(t[c] * area / 3.0, 3))
    return f


def clamp(nodes, predicate):
    """All three DOFs fixed (value 0) at the nodes satisfying ``predicate``."""
    fixed = np.repeat(predicate(nodes), 3)
    return fixed, np.zeros(3 * len(nodes))


# ---------------------------------------------------------------- workloads
def cantilever_tet4(n=(20, 4, 4), size=(100.0, 20.0, 20.0), young=210000.0, nu=0.3,
                    traction=(0.0, 0.0, -10.0)):
    """Box cantilever in Tet4, clamped at x=0, uniform traction on the x=L face."""
    nodes, tets = box_grid(n, size)
    return _cantilever("cantilever_tet4", nodes, tets, size, young, nu, traction, n)


def cantilever_tet10(n=(60, 12, 12), size=(300.0, 60.0, 60.0), young=70000.0, nu=0.33,
                     traction=(0.0, 5.0, -10.0)):
    """Box cantilever in Tet10, clamped at x=0, uniform traction on the x=L face."""
    nodes, tets = box_grid(n, size)
    nodes, conn = to_quadratic_tet(nodes, tets)
    return _cantilever("cantilever_tet10", nodes, conn, size, young, nu, traction, n)


def _cantilever(name, nodes, conn, size, young, nu, traction, n):
    length = float(size[0])
    fixed, imposed = clamp(nodes, lambda x: np.abs(x[:, 0]) < 1e-12)
    f = traction_loads(nodes, conn, lambda x: np.abs(x[:, 0] - length) < 1e-12, traction,
                       3 * len(nodes))
    ne = len(conn)
    return Problem(name, nodes, conn, np.full(ne, float(young)), np.full(ne, float(nu)),
                   fixed, imposed, f, [], {"cells": tuple(n), "size": tuple(size),
                                            "traction": tuple(traction)})


def reference_cantilever(n=(10, 2, 2)):
    """tests/iterative.rs:46-66 — Tet10 cantilever 10x1x1 clamped at x=0, a load Uz = 1 at
    every node of x=10 (the reference's iterative-vs-direct test case)."""
    nodes, tets = box_grid(n, (10.0, 1.0, 1.0))
    nodes, conn = to_quadratic_tet(nodes, tets)
    fixed, imposed = clamp(nodes, lambda x: np.abs(x[:, 0]) < 1e-12)
    f = np.zeros(3 * len(nodes))
    tip = np.flatnonzero(np.abs(nodes[:, 0] - 10.0) < 1e-12)
    f[3 * tip + 2] += 1.0
    ne = len(conn)
    return Problem("reference_cantilever", nodes, conn, np.full(ne, 210e3), np.full(ne, 0.3),
                   fixed, imposed, f, [], {"cells": tuple(n)})


def jittered_cube(cells=100, jitter=0.2, seed_nodes=1, seed_young=2, median=200000.0,
                  sigma=0.5, nu=0.3, pressure=1.0):
    """Tet4 unit cube of cells^3 Kuhn cubes; interior nodes moved uniformly by up to
    ``jitter``*h per axis (SplitMix64 seed ``seed_nodes``, draws x, y, z per interior node in
    index order); per-element E = median*exp(sigma*N(0,1)) (seed ``seed_young``); clamped at
    z=0; pressure on the z=1 faces through face_pressure (positive compresses)."""
    nodes, tets = box_grid((cells, cells, cells), (1.0, 1.0, 1.0))
    h = 1.0 / cells
    m = cells + 1
    k, j, i = np.unravel_index(np.arange(m ** 3), (m, m, m))
    interior = np.flatnonzero((i > 0) & (i < cells) & (j > 0) & (j < cells) & (k > 0)
                              & (k < cells))
    u = uniform(seed_nodes, 3 * len(interior)).reshape(-1, 3)
    nodes[interior] += (2.0 * u - 1.0) * (jitter * h)
    bad = check_tetrahedra(nodes, tets)
    if bad >= 0:
        raise ValueError(f"jittered mesh has an invalid tetrahedron: {bad}")
    ne = len(tets)
    young = median * np.exp(sigma * normal(seed_young, ne))
    fixed, imposed = clamp(nodes, lambda x: np.abs(x[:, 2]) < 1e-12)
    e, fc = faces_on(nodes, tets, lambda x: np.abs(x[:, 2] - 1.0) < 1e-12)
    pressures = list(zip(e.tolist(), fc.tolist(), [float(pressure)] * len(e)))
    return Problem("jittered_cube", nodes, tets, young, np.full(ne, float(nu)), fixed, imposed,
                   np.zeros(3 * len(nodes)), pressures,
                   {"cells": cells, "jitter": jitter, "seeds": (seed_nodes, seed_young)})


def simp_box(n=(80, 40, 40), size=(200.0, 100.0, 100.0), seed=5, rho_min=0.3, e0=210000.0,
             nu=0.3, penal=3.0, load=-1000.0):
    """SIMP compliance problem: Tet4 box, densities rho_e ~ U[rho_min, 1) (seed), E_e =
    e0*rho_e^penal, clamped at x=0, total ``load`` in z spread equally over the x=L nodes
    (as tests/iterative.rs:247-256 spreads its tip load)."""
    nodes, tets = box_grid(n, size)
    ne = len(tets)
    rho = rho_min + (1.0 - rho_min) * uniform(seed, ne)
    fixed, imposed = clamp(nodes, lambda x: np.abs(x[:, 0]) < 1e-12)
    tip = np.flatnonzero(np.abs(nodes[:, 0] - size[0]) < 1e-12)
    f = np.zeros(3 * len(nodes))
    f[3 * tip + 2] += load / len(tip)
    return Problem("simp_box", nodes, tets, e0 * rho ** penal, np.full(ne, float(nu)), fixed,
                   imposed, f, [], {"cells": tuple(n), "rho": rho, "e0": e0, "penal": penal})


def problem_pressures_arrays(problem: Problem):
    """(element, face, value) arrays of the problem's dead pressures."""
    if not problem.pressures:
        z = np.zeros(0, dtype=np.int64)
        return z, z, np.zeros(0)
    p = np.array(problem.pressures, dtype=np.float64)
    return p[:, 0].astype(np.int64), p[:, 1].astype(np.int64), p[:, 2].copy()


# ---------------------------------------------------------------- batched workloads
def face_traction_loads(nodes, conn, faces, tractions, ndof):
    """Consistent nodal loads of a traction vector given per face ((element, face) pairs of
    TET_FACES): A/3 of it at each vertex (Tet4) or mid-node (Tet10), as traction_loads."""
    quadratic = conn.shape[1] == 10
    f = np.zeros(ndof)
    e, fc = faces
    if len(e) == 0:
        return f
    verts = conn[e[:, None], TET_FACES[fc]]  # (nf, 3)
    p = nodes[verts]
    cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    area = np.sqrt((cr ** 2).sum(axis=1)) / 2.0
    if quadratic:
        edge = {tuple(sorted(map(int, ab))): 4 + k for k, ab in enumerate(TET10_EDGES)}
        mids = np.array([[edge[tuple(sorted((int(f3[a]), int(f3[(a + 1) % 3]))))] for a in range(3)]
                         for f3 in TET_FACES])
        carriers = conn[e[:, None], mids[fc]]
    else:
        carriers = verts
    for c in range(3):
        np.add.at(f, 3 * carriers.ravel() + c, np.repeat(tractions[:, c] * area / 3.0, 3))
    return f


def beam_load_cases(n=(40, 8, 8), size=(200.0, 40.0, 40.0), cases=256, seed=3, young=70000.0,
                    nu=0.33):
    """Tet10 beam of n cells (10:2:2 aspect), clamped at x=0. Load case k: on every face of the
    top surface z = H a traction vector of standard-normal components (SplitMix64 seed
    ``seed + k``, three draws per face in face order) as consistent nodal loads. Returns the
    problem (loads of case 0) and the (cases, 3N) load matrix."""
    nodes, tets = box_grid(n, size)
    nodes, conn = to_quadratic_tet(nodes, tets)
    fixed, imposed = clamp(nodes, lambda x: np.abs(x[:, 0]) < 1e-12)
    top = faces_on(nodes, conn[:, :4], lambda x: np.abs(x[:, 2] - size[2]) < 1e-12)
    ndof = 3 * len(nodes)
    F = np.stack([face_traction_loads(nodes, conn, top,
                                      normal(seed + k, 3 * len(top[0])).reshape(-1, 3), ndof)
                  for k in range(cases)])
    ne = len(conn)
    p = Problem("beam_load_cases", nodes, conn, np.full(ne, young), np.full(ne, nu), fixed,
                imposed, F[0].copy(), [], {"cells": tuple(n), "cases": cases})
    return p, F


def box_batch(count=64, seed=4, lo=500.0, hi=20000.0):
    """Tet4 boxes of uneven sizes: N ~ log-uniform in [lo, hi] DOF (seed), cells (nx, m, m) with
    m = ceil((N/48)^(1/3)) and nx = round(N/(3(m+1)^2)) - 1, unit cells; E ~ U[5e4, 2.5e5] and
    nu ~ U[0.2, 0.4] per box (seeds seed+1, seed+2); clamped at x=0, uniform traction
    (0, 0, -1) on the x = nx face."""
    u = uniform(seed, count)
    dofs = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * u)
    young = 5e4 + 2e5 * uniform(seed + 1, count)
    nu = 0.2 + 0.2 * uniform(seed + 2, count)
    out = []
    for k in range(count):
        m = int(np.ceil((dofs[k] / 48.0) ** (1.0 / 3.0)))
        nx = max(1, int(round(dofs[k] / (3.0 * (m + 1) ** 2))) - 1)
        size = (float(nx), float(m), float(m))
        nodes, tets = box_grid((nx, m, m), size)
        p = _cantilever(f"box_{k}", nodes, tets, size, float(young[k]), float(nu[k]),
                        (0.0, 0.0, -1.0), (nx, m, m))
        out.append(p)
    return out


def plastic_bar(n=(120, 24, 24), size=(600.0, 120.0, 120.0), tip=7.0, young=210000.0, nu=0.3):
    """A Tet4 bar clamped at x = 0 (all DOFs) and bent by an imposed tip displacement u_z = tip
    on the x = Lx face (about three times the first-yield deflection 2 sy L^2 / (3 E h) of a
    steel bar with sy = 250 MPa): a plastic zone grows from the clamp as the amplitude ramps up.
    Elastic constants per element; the plasticity parameters are the caller's."""
    nodes, conn = box_grid(n, size)
    nn = len(nodes)
    fixed = np.zeros(3 * nn, dtype=bool)
    imposed = np.zeros(3 * nn)
    x0 = np.abs(nodes[:, 0]) < 1e-9 * size[0]
    x1 = np.abs(nodes[:, 0] - size[0]) < 1e-9 * size[0]
    for c in range(3):
        fixed[3 * np.nonzero(x0)[0] + c] = True
    fixed[3 * np.nonzero(x1)[0] + 2] = True
    imposed[3 * np.nonzero(x1)[0] + 2] = tip
    ne = len(conn)
    return Problem(f"plastic_bar_{n[0]}x{n[1]}x{n[2]}", nodes, conn, np.full(ne, young),
                   np.full(ne, nu), fixed, imposed, np.zeros(3 * nn))
