"""Gmsh 4.1 test inputs: the fixture files of the reference's reader tests (crates/maille-io/
tests/msh.rs), built here, and a writer of generated meshes (Gmsh numbering, reversed elements,
shuffled tags, several blocks, point/line/surface/volume groups) whose expected reading is
known by construction."""
from __future__ import annotations

import numpy as np

from maille_amd import meshes as M

# tests/msh.rs:277-281 — one triangle in a "plaque" surface group
TRIANGLE = ("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n"
            "$PhysicalNames\n1\n2 1 \"plaque\"\n$EndPhysicalNames\n"
            "$Entities\n0 0 1 0\n1 0 0 0 1 1 0 1 1 0\n$EndEntities\n"
            "$Nodes\n1 3 1 3\n2 1 0 3\n1\n2\n3\n0 0 0\n1 0 0\n0 1 0\n$EndNodes\n"
            "$Elements\n1 1 1 1\n2 1 2 1\n1 1 2 3\n$EndElements\n")


def _gmsh_tet10(v):
    """Vertices then Gmsh's mid-nodes (0,1) (1,2) (2,0) (3,0) (3,2) (3,1)."""
    v = [np.array(p, dtype=float) for p in v]
    mids = [(0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)]
    return v + [(v[a] + v[b]) / 2 for a, b in mids]


def _fmt(x):
    return repr(float(x)).replace(".0", "") if float(x).is_integer() else repr(float(x))


def tet10_file():
    """tests/msh.rs:71-96: two Tet10 written Gmsh's way, the second reversed (vertices 0 and 1
    exchanged); groups "coin" (point 1), "base" (triangle 1-2-3), "corps" (both tetrahedra)."""
    a = _gmsh_tet10([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    b = _gmsh_tet10([[3, 0, 0], [2, 0, 0], [2, 1, 0], [2, 0, 1]])
    coords = "".join(f"{_fmt(p[0])} {_fmt(p[1])} {_fmt(p[2])}\n" for p in a + b)
    tags = "".join(f"{i}\n" for i in range(1, 21))
    nodes = lambda first: "".join(f" {i}" for i in range(first, first + 10))
    return ("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n"
            "$PhysicalNames\n3\n0 3 \"coin\"\n2 2 \"base\"\n3 1 \"corps\"\n$EndPhysicalNames\n"
            "$Entities\n1 0 1 1\n1 0 0 0 1 3 0 0 1 0 1\n"
            "1 0 0 0 1 1 0 1 2\n1 0 0 0 1 1 1 1 1 0\n$EndEntities\n"
            f"$Nodes\n1 20 1 20\n3 1 0 20\n{tags}{coords}$EndNodes\n"
            "$Elements\n3 4 1 4\n0 1 15 1\n1 1\n2 1 2 1\n2 1 2 3\n3 1 11 2\n"
            f"3{nodes(1)}\n4{nodes(11)}\n$EndElements\n")


def tet10_file_with_line(line):
    """tests/msh.rs:156-170: tet10_file plus a line group "arete" of one 3-node line."""
    return (tet10_file()
            .replace("$PhysicalNames\n3\n", "$PhysicalNames\n4\n1 4 \"arete\"\n")
            .replace("$Entities\n1 0 1 1\n1 0 0 0 1 3 0 0 1 0 1\n",
                     "$Entities\n1 1 1 1\n1 0 0 0 1 3 0 0 1 0 1\n1 0 0 0 0 0 1 1 4 0\n")
            .replace("$Elements\n3 4 1 4\n0 1 15 1\n1 1\n",
                     f"$Elements\n4 5 1 5\n0 1 15 1\n1 1\n1 1 8 1\n5 {line}\n"))


# --------------------------------------------------------------- generated meshes
GMSH_OF_MAILLE_TET10 = [0, 1, 2, 3, 4, 5, 6, 7, 9, 8]  # the swap is its own inverse


def reverse(cell):
    """A positive cell in Maille numbering made negative the way orient() undoes it."""
    c = list(cell)
    c[0], c[1] = c[1], c[0]
    if len(c) == 10:
        c[5], c[6] = c[6], c[5]
        c[7], c[8] = c[8], c[7]
    return c


def write_msh(nodes, cells, groups, seed=0, blocks=3, flip=0.5, sparse_tags=False):
    """A Gmsh 4.1 file of ``cells`` (Maille numbering, positive) over ``nodes``.

    groups: list of (name, dim, items) with items node indices (dim 0), vertex pairs (dim 1,
    written as 2- or 3-node lines), vertex triples (dim 2, written as 3- or 6-node triangles)
    or None for the volume group (dim 3, every cell). Node tags are a shuffled (or sparse)
    numbering; cells are split over ``blocks`` volume entities; a fraction ``flip`` of the cells
    is written reversed. The reading must give back ``cells`` and ``nodes`` exactly."""
    rng = np.random.default_rng(seed)
    n, (ne, nen) = len(nodes), cells.shape
    if sparse_tags:
        tags = np.sort(rng.choice(np.arange(1, 50 * n + 10**9, 997), n, replace=False))
        tags = rng.permutation(tags)
    else:
        tags = rng.permutation(n) + 1
    quad = nen == 10
    mid = {}
    if quad:
        for c in cells:
            for k, (a, b) in enumerate(M.TET10_EDGES):
                mid[frozenset((int(c[a]), int(c[b])))] = int(c[4 + k])
    out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat"]
    out += ["$PhysicalNames", str(len(groups))]
    for g, (name, dim, _) in enumerate(groups):
        out.append(f'{dim} {g + 1} "{name}"')
    out.append("$EndPhysicalNames")
    # entities: one per group of dimension < 3 (tag = group index + 1), ``blocks`` volumes
    ent = {0: [], 1: [], 2: [], 3: []}
    for g, (name, dim, _) in enumerate(groups):
        if dim < 3:
            ent[dim].append((g + 1, [g + 1]))
    vol_groups = [g + 1 for g, (_, dim, _) in enumerate(groups) if dim == 3]
    for v in range(blocks):
        ent[3].append((v + 1, vol_groups))
    out += ["$Entities", " ".join(str(len(ent[d])) for d in range(4))]
    for d in range(4):
        for tag, phys in ent[d]:
            box = "0 0 0" if d == 0 else "0 0 0 1 1 1"
            tail = "" if d == 0 else " 0"
            out.append(f"{tag} {box} {len(phys)}" + "".join(f" {p}" for p in phys) + tail)
    out.append("$EndEntities")
    # nodes in index order (the reading numbers nodes by their place in the file), two blocks
    halves = np.array_split(np.arange(n), 2)
    out += ["$Nodes", f"{len(halves)} {n} {int(tags.min())} {int(tags.max())}"]
    for h, part in enumerate(halves):
        out.append(f"3 {h + 1} 0 {len(part)}")
        out += [str(int(tags[i])) for i in part]
        out += [" ".join(repr(float(v)) for v in nodes[i]) for i in part]
    out.append("$EndNodes")
    # elements: lower-dimensional groups first, then the cells over ``blocks`` blocks
    eb, count, etag = [], 0, 1
    for g, (name, dim, items) in enumerate(groups):
        if dim == 3:
            continue
        if dim == 0:
            lines = [str(int(tags[i])) for i in items]
            t = 15
        elif dim == 1:
            t = 8 if quad else 1
            lines = []
            for a, b in items:
                ids = [a, b] + ([mid[frozenset((a, b))]] if quad else [])
                lines.append(" ".join(str(int(tags[i])) for i in ids))
        else:
            t = 9 if quad else 2
            lines = []
            for a, b, c in items:
                ids = [a, b, c]
                if quad:
                    ids += [mid[frozenset((a, b))], mid[frozenset((b, c))], mid[frozenset((c, a))]]
                lines.append(" ".join(str(int(tags[i])) for i in ids))
        eb.append((dim, g + 1, t, lines))
    parts = np.array_split(np.arange(ne), blocks)
    for v, part in enumerate(parts):
        lines = []
        for e in part:
            c = list(map(int, cells[e]))
            if rng.random() < flip:
                c = reverse(c)
            if quad:
                c = [c[k] for k in GMSH_OF_MAILLE_TET10]
            lines.append(" ".join(str(int(tags[i])) for i in c))
        eb.append((3, v + 1, 11 if quad else 4, lines))
    total = sum(len(b[3]) for b in eb)
    out += ["$Elements", f"{len(eb)} {total} 1 {total}"]
    for dim, tag, t, lines in eb:
        out.append(f"{dim} {tag} {t} {len(lines)}")
        for ln in lines:
            out.append(f"{etag} {ln}")
            etag += 1
    out.append("$EndElements")
    return "\n".join(out) + "\n"


def first_owner(cells, positions, key_len, wanted):
    """Brute force: for each wanted vertex tuple, the first (element, local index) whose
    positions' first ``key_len`` vertices make the same set."""
    out = {}
    want = {frozenset(w) for w in wanted}
    for e, c in enumerate(cells):
        for k, pos in enumerate(positions):
            key = frozenset(int(c[p]) for p in pos[:key_len])
            if key in want and key not in out:
                out[key] = (e, k)
    return out
