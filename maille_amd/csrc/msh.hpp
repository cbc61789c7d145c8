// Gmsh 4.1 ASCII reader (maille-io/src/msh/reader.rs, numbering.rs): nodes, cells in Maille's
// numbering with positive orientation, and the physical groups as named sets.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ml {
namespace msh {

// Refusal kinds, in the order of ml_msh_error (include/maille_hip.h).
enum ErrorKind : int32_t {
  E_SYNTAX = 1,
  E_MISSING_SECTION = 2,
  E_UNSUPPORTED_VERSION = 3,
  E_BINARY = 4,
  E_UNSUPPORTED_ELEMENT = 5,
  E_NO_CELLS = 6,
  E_MIXED_CELL_KINDS = 7,
  E_NON_PLANAR = 8,
  E_DUPLICATE_SET_NAME = 9,
  E_DUPLICATE_NODE_TAG = 10,
  E_UNKNOWN_NODE = 11,
  E_BOUNDARY_FACE_NOT_FOUND = 12,
  E_BOUNDARY_EDGE_NOT_FOUND = 13,
};

struct Error : std::runtime_error {
  int32_t kind;
  int64_t line;  // 1-based, 0 when the refusal has no line
  Error(int32_t k, int64_t l, const std::string& m) : std::runtime_error(m), kind(k), line(l) {}
};

// mod.rs:97-109: sorted, duplicate-free; faces and edges as (element, local index)
struct Set {
  std::vector<int64_t> nodes;
  std::vector<std::pair<int64_t, int64_t>> faces;
  std::vector<std::pair<int64_t, int64_t>> edges;
  std::vector<int64_t> elements;
};

// mod.rs:111-120. nen is 3 (Tri3), 6 (Tri6), 4 (Tet4) or 10 (Tet10); cells is ncells x nen.
struct File {
  std::vector<double> xyz;
  int32_t nen = 0;
  int32_t dim = 0;
  std::vector<int64_t> cells;
  std::map<std::string, Set> sets;  // byte order of the names, as a BTreeMap<String, _>
  int64_t n_nodes() const { return (int64_t)xyz.size() / 3; }
  int64_t n_cells() const { return nen ? (int64_t)cells.size() / nen : 0; }
};

// reader.rs:23-39. Throws Error for a refused file.
File read(const char* text, size_t length);

}  // namespace msh
}  // namespace ml
