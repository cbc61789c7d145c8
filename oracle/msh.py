"""The oracle's Gmsh 4.1 ASCII reader: test infrastructure, not part of the product.

A plain-Python restatement of maille-io's ``.msh`` reader, line by line where the reference
is (crates/maille-io/src/msh/reader.rs, numbering.rs, mod.rs of timothee-faget/maille
v0.14.0), used by ``tests/`` as the checker of the library's native reader. It builds the
reference's full face and edge maps (first element wins) rather than the library's query-only
lookup, so the two reach the same sets by different routes.

Pinned by: the reference's own reader tests restated in tests/test_msh.py (tests/msh.rs: Tet10
numbering and orientation, sets, line groups, the refusals and their line numbers), the
reference's example file examples/cube.msh (six Tet4, three of them written with negative
volume), and invariants (positive volumes, mid-nodes at edge midpoints, quadratic fields
reproduced by the Tet10 interpolation).

Deliberate narrowing: whitespace is ASCII whitespace (Rust's ``trim``/``split_whitespace``
also accept the Unicode spaces), which no Gmsh file contains.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

# shape.rs:129-135 — the tetrahedron's faces and the Tet10 edge order (mid-node 4 + k on edge k)
TET_FACES = ((0, 1, 2), (0, 3, 1), (1, 3, 2), (2, 3, 0))
TET10_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))
TRI6_EDGES = ((0, 1), (1, 2), (2, 0))

KINDS = {2: "tri3", 9: "tri6", 4: "tet4", 11: "tet10"}  # mod.rs:80-88
NODE_COUNT = {"tri3": 3, "tri6": 6, "tet4": 4, "tet10": 10}
DIM = {"tri3": 2, "tri6": 2, "tet4": 3, "tet10": 3}
# reader.rs:250-261 — Gmsh types read: (dimension, nodes per element)
ELEMENT_INFO = {15: (0, 1), 1: (1, 2), 8: (1, 3), 2: (2, 3), 9: (2, 6), 4: (3, 4), 11: (3, 10)}

_FLOAT = re.compile(
    r"[+-]?(?:(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?|(?i:inf|infinity|nan))\Z")
_INT = re.compile(r"[+-]?[0-9]+\Z")
_UINT = re.compile(r"\+?[0-9]+\Z")
_WS = " \t\n\r\x0b\x0c"
_SPLIT = re.compile("[ \t\n\r\x0b\x0c]+")


def _split(s):
    """split_whitespace over ASCII whitespace."""
    return [t for t in _SPLIT.split(s) if t]


class MshError(ValueError):
    """IoError's reader variants (error.rs:9-65): ``kind`` names the variant, ``line`` is the
    1-based line or None."""

    def __init__(self, kind, line=None, message="", **extra):
        super().__init__(f"line {line}: {message}" if line is not None else message)
        self.kind, self.line = kind, line
        for k, v in extra.items():
            setattr(self, k, v)


@dataclass
class MshSet:
    """mod.rs:97-109 — sorted, duplicate-free node, face, edge and element lists."""
    nodes: list = field(default_factory=list)
    faces: list = field(default_factory=list)
    edges: list = field(default_factory=list)
    elements: list = field(default_factory=list)


@dataclass
class MshFile:
    """mod.rs:111-120 — nodes as (x, y, z) tuples, cells in Maille numbering."""
    nodes: list
    kind: str
    cells: list
    sets: dict


def _syntax(line, message):
    return MshError("syntax", line, message)


def _split_sections(text):
    """reader.rs:46-80."""
    sections, current, last = {}, None, 0
    for i, raw in enumerate(text.split("\n")):
        n, line = i + 1, raw.strip(_WS)
        if not line:
            continue
        last = n
        dollar = line[1:] if line.startswith("$") else None
        if current is None:
            if dollar is not None and not dollar.startswith("End"):
                current = (dollar, [])
            continue
        name, body = current
        if dollar is not None and dollar == "End" + name:
            sections[name] = body
            current = None
        elif dollar is not None:
            raise _syntax(n, f"'${dollar}' inside section ${name}: $End{name} is missing")
        else:
            body.append((n, line))
    if current is not None:
        raise _syntax(last, f"section ${current[0]} is not closed: $End{current[0]} is missing")
    return sections


class _Cursor:
    """reader.rs:83-114."""

    def __init__(self, name, body):
        self.name, self.lines, self.pos = name, body, 0

    def next_line(self):
        if self.pos >= len(self.lines):
            at = self.lines[-1][0] if self.lines else 0
            raise _syntax(at, f"section ${self.name} ends too early")
        self.pos += 1
        return self.lines[self.pos - 1]

    def next(self):
        n, line = self.next_line()
        return n, _split(line)


def _num(line, toks, i, what, kind):
    """reader.rs:116-126 with Rust's FromStr grammars for f64, i64 and usize/u32."""
    if i >= len(toks):
        raise _syntax(line, f"{what} is missing")
    tok = toks[i]
    pattern = {"f": _FLOAT, "i": _INT, "u": _UINT}[kind]
    if not pattern.match(tok):
        raise _syntax(line, f"{what}: '{tok}' is not a valid number")
    if kind == "f":
        return float(tok)
    v = int(tok)
    if (kind == "i" and not -2**63 <= v < 2**63) or (kind == "u" and v >= 2**64):
        raise _syntax(line, f"{what}: '{tok}' is not a valid number")
    return v


def _check_format(body):
    """reader.rs:128-142."""
    line, toks = _Cursor("MeshFormat", body).next()
    version = toks[0] if toks else ""
    if version != "4.1":
        raise MshError("unsupported_version", line,
                       f"format {version} is not supported; export with Mesh.MshFileVersion = 4.1",
                       found=version)
    if len(toks) < 2 or toks[1] != "0":
        raise MshError("binary", line, "binary file; export with Mesh.Binary = 0")


def _parse_names(body):
    """reader.rs:144-169."""
    c = _Cursor("PhysicalNames", body)
    line, toks = c.next()
    count = _num(line, toks, 0, "the number of physical groups", "u")
    names = {}
    for _ in range(count):
        line, text = c.next_line()
        op, cl = text.find('"'), text.rfind('"')
        if op < 0 or cl <= op:
            raise _syntax(line, "the group name must be in double quotes")
        head = _split(text[:op])
        dim = _num(line, head, 0, "the group dimension", "u")
        tag = _num(line, head, 1, "the group number", "i")
        names[(dim, tag)] = text[op + 1:cl]
    return names


def _parse_entities(body):
    """reader.rs:172-199."""
    c = _Cursor("Entities", body)
    line, toks = c.next()
    counts = [_num(line, toks, d, "the number of entities", "u") for d in range(4)]
    entities = {}
    for dim, n in enumerate(counts):
        for _ in range(n):
            line, toks = c.next()
            tag = _num(line, toks, 0, "the entity number", "i")
            at = 4 if dim == 0 else 7
            count = _num(line, toks, at, "the entity's number of physical groups", "u")
            entities[(dim, tag)] = [_num(line, toks, at + 1 + k, "a physical group number", "i")
                                    for k in range(count)]
    return entities


def _parse_nodes(body):
    """reader.rs:202-247."""
    c = _Cursor("Nodes", body)
    line, toks = c.next()
    blocks = _num(line, toks, 0, "the number of node blocks", "u")
    total = _num(line, toks, 1, "the number of nodes", "u")
    nodes, index = [], {}
    for _ in range(blocks):
        bl, btoks = c.next()
        count = _num(bl, btoks, 3, "the number of nodes of the block", "u")
        tags = []
        for _ in range(count):
            tl, ttoks = c.next()
            tags.append((tl, _num(tl, ttoks, 0, "the node number", "i")))
        for tag_line, tag in tags:
            cl, ctoks = c.next()
            p = tuple(_num(cl, ctoks, k, "xyz"[k], "f") for k in range(3))
            if not all(math.isfinite(v) for v in p):
                raise _syntax(cl, f"a coordinate of node {tag} is not finite (NaN or infinite)")
            if tag in index:
                raise MshError("duplicate_node_tag", tag_line, f"node {tag} is defined twice",
                               tag=tag)
            index[tag] = len(nodes)
            nodes.append(p)
    if len(nodes) != total:
        raise _syntax(line, f"{total} nodes announced, {len(nodes)} found")
    return nodes, index


def _parse_elements(body, index):
    """reader.rs:273-317 — blocks as (entity, gmsh type, dimension, [(line, node indices)])."""
    c = _Cursor("Elements", body)
    line, toks = c.next()
    nblocks = _num(line, toks, 0, "the number of element blocks", "u")
    out = []
    for _ in range(nblocks):
        bl, t = c.next()
        edim = _num(bl, t, 0, "the entity dimension", "u")
        etag = _num(bl, t, 1, "the entity number", "i")
        gtype = _num(bl, t, 2, "the element type", "u")
        if gtype >= 2**32:
            raise _syntax(bl, f"the element type: '{t[2]}' is not a valid number")
        count = _num(bl, t, 3, "the number of elements of the block", "u")
        if gtype not in ELEMENT_INFO:
            raise MshError("unsupported_element", bl,
                           f"element type {gtype} ({gmsh_name(gtype)}) is not supported",
                           gmsh_type=gtype)
        dim, width = ELEMENT_INFO[gtype]
        elements = []
        for _ in range(count):
            el, et = c.next()
            if len(et) != width + 1:
                raise _syntax(el, f"an element of type {gtype} must have {width} nodes, "
                                  f"it has {max(len(et) - 1, 0)}")
            ids = []
            for k in range(1, width + 1):
                tag = _num(el, et, k, "a node number", "i")
                if tag not in index:
                    raise MshError("unknown_node", el, f"node {tag} is not defined in $Nodes",
                                   tag=tag)
                ids.append(index[tag])
            elements.append((el, ids))
        out.append(((edim, etag), gtype, dim, elements))
    return out


def gmsh_name(t):
    """A readable name of a Gmsh element type for the refusal message."""
    return {3: "quadrangle", 5: "hexahedron", 6: "prism", 7: "pyramid", 10: "quadrangle-9",
            12: "hexahedron-27", 16: "quadrangle-8", 17: "hexahedron-20"}.get(t, f"type {t}")


def swap_tet10_mids(nodes):
    """numbering.rs:11-15 — Gmsh's mid-nodes 8 = (3,2), 9 = (3,1) become Maille's 8 = (1,3),
    9 = (2,3)."""
    if len(nodes) == 10:
        nodes[8], nodes[9] = nodes[9], nodes[8]


def orient(kind, nodes, points):
    """numbering.rs:20-46 — swap vertices 0 and 1 (and the mid-nodes that follow them) of a
    clockwise triangle or a negative tetrahedron."""
    p = lambda i: points[nodes[i]]
    if kind in ("tri3", "tri6"):
        reversed_ = ((p(1)[0] - p(0)[0]) * (p(2)[1] - p(0)[1])
                     - (p(2)[0] - p(0)[0]) * (p(1)[1] - p(0)[1])) < 0.0
    else:
        e = lambda k: [p(k)[c] - p(0)[c] for c in range(3)]
        a, b, c = e(1), e(2), e(3)
        det = (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0])
               + a[2] * (b[0] * c[1] - b[1] * c[0]))
        reversed_ = det < 0.0
    if not reversed_:
        return
    nodes[0], nodes[1] = nodes[1], nodes[0]
    if kind == "tri6":
        nodes[4], nodes[5] = nodes[5], nodes[4]
    elif kind == "tet10":
        nodes[5], nodes[6] = nodes[6], nodes[5]
        nodes[7], nodes[8] = nodes[8], nodes[7]


def face_count(kind):
    return 3 if DIM[kind] == 2 else 4


def face_positions(kind, f):
    """numbering.rs:59-72."""
    if kind == "tri3":
        return list(TRI6_EDGES[f])
    if kind == "tri6":
        return [*TRI6_EDGES[f], 3 + f]
    a, b, c = TET_FACES[f]
    if kind == "tet4":
        return [a, b, c]

    def mid(x, y):
        for k, e in enumerate(TET10_EDGES):
            if e in ((x, y), (y, x)):
                return 4 + k
        return 4

    return [a, b, c, mid(a, b), mid(b, c), mid(c, a)]


def edge_positions(kind, k):
    """numbering.rs:85-91."""
    if DIM[kind] == 2:
        return face_positions(kind, k)
    a, b = TET10_EDGES[k]
    return [a, b] if kind == "tet4" else [a, b, 4 + k]


def face_key(vertices):
    """numbering.rs:94-99 (the padding of a 2-vertex key is implicit in the tuple length)."""
    return tuple(sorted(vertices))


def _check_planar(points):
    """reader.rs:461-471."""
    span = 0.0
    for p in points:
        span = max(span, abs(p[0]), abs(p[1]))
    tol = 1e-9 * max(span, 2.2250738585072014e-308)
    for p in points:
        if abs(p[2]) > tol:
            raise MshError("non_planar", None, f"triangle mesh outside the plane z = 0 "
                                               f"(z = {p[2]})", z=p[2])


def _assemble(points, blocks, names, entities):
    """reader.rs:320-458."""
    types = {b[1] for b in blocks}
    if types & {4, 11}:
        dim = 3
    elif types & {2, 9}:
        dim = 2
    else:
        raise MshError("no_cells", None, "the file holds no triangle and no tetrahedron")
    if dim == 2:
        _check_planar(points)
    kind, cells, entity_cells = None, [], {}
    for entity, gtype, bdim, elements in blocks:
        if bdim != dim:
            continue
        k = KINDS[gtype]
        if kind is not None and kind != k:
            raise MshError("mixed_cell_kinds", None, f"the file mixes {kind} and {k} elements",
                           first=kind, found=k)
        kind = k
        for _, ids in elements:
            ids = list(ids)
            if k == "tet10":
                swap_tet10_mids(ids)
            orient(k, ids, points)
            entity_cells.setdefault(entity, []).append(len(cells))
            cells.append(ids)
    faces = {}
    for e, cell in enumerate(cells):
        for f in range(face_count(kind)):
            key = face_key([cell[p] for p in face_positions(kind, f)[:dim]])
            faces.setdefault(key, (e, f))
    edges = {}
    if dim == 3:
        for e, cell in enumerate(cells):
            for k in range(6):
                a, b = edge_positions(kind, k)[:2]
                edges.setdefault(face_key([cell[a], cell[b]]), (e, k))
    sets, owner = {}, {}
    for entity, gtype, bdim, elements in blocks:
        physical = entities.get(entity)
        if physical is None:
            continue
        for tag in physical:
            group = (entity[0], tag)
            name = names.get(group, str(tag))
            if owner.setdefault(name, group) != group:
                raise MshError("duplicate_set_name", None,
                               f"two physical groups are named '{name}'", name=name)
            s = sets.setdefault(name, MshSet())
            if bdim == dim:
                s.elements.extend(entity_cells.get(entity, []))
            elif bdim + 1 == dim:
                for line, ids in elements:
                    hit = faces.get(face_key(ids[:dim]))
                    if hit is None:
                        raise MshError("boundary_face_not_found", line,
                                       f"an element of group '{name}' is no element face",
                                       set=name)
                    s.faces.append(hit)
                    s.nodes.extend(cells[hit[0]][p] for p in face_positions(kind, hit[1]))
            elif bdim == 1:
                for line, ids in elements:
                    hit = edges.get(face_key(ids[:2]))
                    if hit is None:
                        raise MshError("boundary_edge_not_found", line,
                                       f"a line of group '{name}' is no element edge", set=name)
                    s.edges.append(hit)
                    s.nodes.extend(cells[hit[0]][p] for p in edge_positions(kind, hit[1]))
            else:
                s.nodes.extend(i for _, ids in elements for i in ids)
    for s in sets.values():
        s.nodes = sorted(set(s.nodes))
        s.faces = sorted(set(s.faces))
        s.edges = sorted(set(s.edges))
        s.elements = sorted(set(s.elements))
    return MshFile(points, kind, cells, dict(sorted(sets.items(), key=lambda kv: kv[0].encode())))


def read(text: str) -> MshFile:
    """reader.rs:23-39."""
    sections = _split_sections(text)

    def section(name):
        if name not in sections:
            raise MshError("missing_section", None, f"section ${name} is missing", name=name)
        return sections[name]

    _check_format(section("MeshFormat"))
    names = _parse_names(sections["PhysicalNames"]) if "PhysicalNames" in sections else {}
    entities = _parse_entities(sections["Entities"]) if "Entities" in sections else {}
    points, index = _parse_nodes(section("Nodes"))
    blocks = _parse_elements(section("Elements"), index)
    return _assemble(points, blocks, names, entities)
