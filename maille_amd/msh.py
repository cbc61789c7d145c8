"""Gmsh 4.1 ASCII input (crates/maille-io/src/msh of timothee-faget/maille v0.14.0).

``read(text)`` / ``read_file(path)`` return an ``MshFile``: the nodes, the cells in Maille's
numbering (Tet10 mid-nodes 8 and 9 swapped from Gmsh's order, numbering.rs:11-15) with positive
orientation (numbering.rs:20-46), and the physical groups as named sets (mod.rs:91-109) of
nodes, faces (element, face), edges (element, edge) and elements. The parse runs in the native
library (``ml_msh_read``); a refused file raises ``MshError`` with the reference's variant as
``kind`` and the 1-based ``line`` (None for the variants without one), as IoError does
(error.rs:9-65).
"""
from __future__ import annotations

import ctypes as C
import mmap
import os
from dataclasses import dataclass, field

import numpy as np

from . import ML_INVALID_FILE, _check, lib

ERROR_KINDS = {1: "syntax", 2: "missing_section", 3: "unsupported_version", 4: "binary",
               5: "unsupported_element", 6: "no_cells", 7: "mixed_cell_kinds",
               8: "non_planar", 9: "duplicate_set_name", 10: "duplicate_node_tag",
               11: "unknown_node", 12: "boundary_face_not_found", 13: "boundary_edge_not_found"}
KINDS = {3: "tri3", 6: "tri6", 4: "tet4", 10: "tet10"}
# variants without a line (error.rs:24-47)
_LINELESS = {"missing_section", "no_cells", "mixed_cell_kinds", "non_planar",
             "duplicate_set_name"}


class MshError(ValueError):
    """A refused ``.msh`` file (IoError's reader variants): ``kind`` and ``line``."""

    def __init__(self, kind, line, message):
        super().__init__(message)
        self.kind, self.line = kind, line


def _empty_pairs():
    return np.zeros((0, 2), dtype=np.int64)


@dataclass
class MshSet:
    """Set (mod.rs:97-109): sorted and duplicate-free."""
    nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    faces: np.ndarray = field(default_factory=_empty_pairs)
    edges: np.ndarray = field(default_factory=_empty_pairs)
    elements: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))


@dataclass
class MshFile:
    """MshFile (mod.rs:111-120): nodes (N, 3), kind, cells (ne, nen) in Maille numbering, sets
    by name in byte order of the names."""
    nodes: np.ndarray
    kind: str
    cells: np.ndarray
    sets: dict

    @property
    def dim(self) -> int:
        return 2 if self.kind in ("tri3", "tri6") else 3


def _read_buffer(address: int, length: int) -> MshFile:
    L = lib()
    h = C.c_void_p()
    kind = C.c_int32()
    line = C.c_int64()
    st = L.ml_msh_read(address, length, C.byref(h), C.byref(kind), C.byref(line))
    if st == ML_INVALID_FILE:
        msg = L.ml_last_error().decode(errors="replace")
        k = ERROR_KINDS.get(kind.value, str(kind.value))
        raise MshError(k, None if k in _LINELESS else line.value, msg)
    _check(st)
    try:
        sizes = np.zeros(5, dtype=np.int64)
        _check(L.ml_msh_sizes(h, sizes.ctypes.data_as(C.POINTER(C.c_int64))))
        n, nen, ne, nsets = (int(v) for v in sizes[:4])
        xyz = np.zeros((n, 3))
        cells = np.zeros((ne, nen), dtype=np.int64)
        _check(L.ml_msh_geometry(h, xyz.ctypes.data_as(C.POINTER(C.c_double)),
                                 cells.ctypes.data_as(C.POINTER(C.c_int64))))
        sets = {}
        counts = np.zeros(5, dtype=np.int64)
        cp = counts.ctypes.data_as(C.POINTER(C.c_int64))
        for k in range(nsets):
            _check(L.ml_msh_set_info(h, k, None, 0, cp))
            buf = C.create_string_buffer(int(counts[4]) + 1)
            _check(L.ml_msh_set_info(h, k, buf, len(buf), cp))
            s = MshSet(np.zeros(counts[0], dtype=np.int64),
                       np.zeros((counts[1], 2), dtype=np.int64),
                       np.zeros((counts[2], 2), dtype=np.int64),
                       np.zeros(counts[3], dtype=np.int64))
            ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
            _check(L.ml_msh_set_items(h, k, ptr(s.nodes), ptr(s.faces), ptr(s.edges),
                                      ptr(s.elements)))
            sets[buf.raw[:counts[4]].decode("utf-8", errors="surrogateescape")] = s
        return MshFile(xyz, KINDS[nen], cells, sets)
    finally:
        L.ml_msh_destroy(h)


def read(text) -> MshFile:
    """msh::read (reader.rs:23-39) of a str or bytes."""
    data = text.encode("utf-8") if isinstance(text, str) else bytes(text)
    buf = C.create_string_buffer(data, len(data)) if data else None
    return _read_buffer(C.addressof(buf) if buf is not None else 0, len(data))


def read_file(path) -> MshFile:
    """Reads a ``.msh`` file through a private memory map (no copy of the text)."""
    size = os.path.getsize(path)
    if size == 0:
        return _read_buffer(0, 0)
    with open(path, "rb") as fh, mmap.mmap(fh.fileno(), 0, access=mmap.ACCESS_COPY) as mm:
        anchor = C.c_char.from_buffer(mm)
        try:
            return _read_buffer(C.addressof(anchor), size)
        finally:
            del anchor
