// Host-side sparsity work of the solve path: the DOF graph, reverse Cuthill-McKee numbering and
// the lower-triangle CSR pattern (bit-exact to maille-core), plus the device layouts derived
// from them (gather lists for assembly, the sliced-ELL symmetric SpMV layout, the transposed
// factor layout and the row blocks of the dataflow triangular solves).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ml {

typedef int64_t i64;

// iterative.rs:16-22: neighbours of v are targets[offsets[v]..offsets[v+1]], in first-seen order
struct Graph {
  std::vector<i64> offsets;
  std::vector<int32_t> targets;
  i64 n() const { return (i64)offsets.size() - 1; }
  i64 degree(i64 v) const { return offsets[v + 1] - offsets[v]; }
};

// iterative.rs:38-76 (Graph::from_elements); elements as CSR lists of vertices < n
Graph graph_from_elements(i64 n, i64 n_elements, const i64* eoff, const i64* eitems);

// solve.rs:200-245 (reverse_cuthill_mckee): order[new] = old
std::vector<i64> reverse_cuthill_mckee(const Graph& g);

// iterative.rs:79-114: lower triangle in the numbering rank, ascending columns, diagonal last
struct LowerPattern {
  std::vector<i64> offsets;
  std::vector<int32_t> cols;
  i64 n() const { return (i64)offsets.size() - 1; }
  i64 nnz() const { return (i64)cols.size(); }
  // Csr::add's binary search (iterative.rs:121-128); -1 when (i, j) is outside the profile
  i64 find(i64 i, i64 j) const;
};
LowerPattern lower_pattern(const Graph& g, const std::vector<i64>& order,
                           const std::vector<i64>& rank);

// Sliced ELL, 64 rows per slice, of the full symmetric matrix; row j holds, in this order,
// the diagonal, the lower entries (ascending column) and the upper entries (ascending row):
// exactly the order in which Csr::multiply (iterative.rs:131-142) accumulates y[j], so a lane
// that sums its row sequentially reproduces the reference bit for bit.
struct SellLayout {
  static constexpr int C = 64;
  i64 n = 0, slices = 0;
  std::vector<i64> slice_offsets;  // slices + 1, in entries
  std::vector<int32_t> cols;       // padding: the row itself
  std::vector<int64_t> src;        // lower-CSR index of each entry, -1 for padding
  i64 entries() const { return (i64)cols.size(); }
};
SellLayout sell_layout(const LowerPattern& lp);

// Rows of L^T (the columns of L): for row j, the rows i > j with l_ij != 0, ascending i, and
// the lower-CSR index of l_ij.
struct TransposeLayout {
  std::vector<i64> offsets;
  std::vector<int32_t> rows;
  std::vector<int64_t> src;
};
TransposeLayout transpose_layout(const LowerPattern& lp);

// DAG depth of the lower-triangular solve (longest dependency chain, in rows).
i64 dag_depth(const LowerPattern& lp);

// Row blocks of the dataflow triangular solve over a strictly-lower layout (off[n+1], col[]):
// consecutive rows, at most `max_rows` per block, and a new block wherever a row's level (its
// longest dependency chain) does not exceed the previous row's. Inside a block the rows then
// form one dependency chain, so solving them one after the other costs no false ordering;
// across blocks the solve only waits for true dependencies. Returns the block starts, n last.
std::vector<int32_t> dataflow_blocks(int64_t n, const int32_t* off, const int32_t* col,
                                     int max_rows);

// One dataflow triangular solve in logical row order (strictly-lower layout off/col, entry
// sources src into the factor's value array, diagonal sources dsrc). Rows are cut into blocks
// as dataflow_blocks does, also capped so a block holds at most xcap entries outside it and
// icap inside it; each row's entries are split into the external part (columns before the
// block) and the internal part (columns inside, stored as the 0..63 offset from the block
// start), both block-contiguous so a wave loads a whole block with coalesced reads.
struct TriLayout {
  std::vector<int32_t> bstart;           // blocks + 1
  std::vector<int32_t> border;           // processing order: blocks sorted by block-DAG level
  std::vector<int32_t> xoff, xcol, xsrc;  // external entries: rows + 1, entries, sources
  std::vector<int32_t> ioff, ik, isrc;    // internal entries
  std::vector<int32_t> dsrc;              // diagonal source per row
  int max_block_ext = 0, max_block_int = 0;
};
TriLayout tri_layout(int64_t n, const int32_t* off, const int32_t* col, const int32_t* src,
                     const int32_t* dsrc, int max_rows, int xcap, int icap, std::string* err);

// Every index array the device kernels of one operator read, built on the host from the lower
// pattern: the 32-bit CSR (IC(0)), the sliced-ELL symmetric layout (SpMV) and the two
// triangular-solve layouts. The backward solve L^T z = y runs in logical rows l = n-1-j: row l
// holds the entries l_ij, i > j, of column j of L in decreasing i, at logical column n-1-i.
struct DeviceLayouts {
  int64_t n = 0, nnz = 0, depth = 0, slices = 0;
  int max_row = 0;  // longest row of L, diagonal included
  std::vector<int32_t> off32, col32;
  std::vector<int32_t> soff, scol, slen, ssrc;
  TriLayout fw, bw;
};
// Returns an empty string or the reason the pattern cannot be laid out (row too long).
std::string device_layouts(const LowerPattern& lp, int xcap, int icap, DeviceLayouts* out);

// Verifies every bound the kernels rely on (index ranges, block capacities, coverage and order
// of each row's entries, topological block order); empty string when all hold.
std::string check_layouts(const LowerPattern& lp, const DeviceLayouts& d, int xcap, int icap);

// The preconditioner z = (L L^T)^-1 r computed on the host through the triangular-solve
// layouts, block by block in processing order, each row summed in the order the device sums
// it. l holds the factor's values in lower-CSR order.
void layout_precondition(const DeviceLayouts& d, const double* l, const double* r, double* z);

}  // namespace ml
