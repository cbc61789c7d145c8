// Host-side sparsity work of the solve path: the DOF graph, reverse Cuthill-McKee numbering and
// the lower-triangle CSR pattern (bit-exact to maille-core), plus the device layouts derived
// from them (gather lists for assembly, the sliced-ELL symmetric SpMV layout, the transposed
// factor layout and the row blocks of the dataflow triangular solves).
#pragma once

#include <cstdint>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace ml {

typedef int64_t i64;

// Threads of the host-side builders whose exceptions reach the caller: run(f) starts f on a
// thread of its own; join() waits for all of them and rethrows the first exception one raised.
// The destructor joins any thread still running (when the caller unwinds before join()).
class ThreadGroup {
 public:
  template <class F>
  void run(F f) {
    threads_.emplace_back([this, f]() mutable {
      try {
        f();
      } catch (...) {
        std::lock_guard<std::mutex> lock(mutex_);
        if (!error_) error_ = std::current_exception();
      }
    });
  }
  void join() {
    for (auto& t : threads_) t.join();
    threads_.clear();
    if (error_) {
      std::exception_ptr e = error_;
      error_ = nullptr;
      std::rethrow_exception(e);
    }
  }
  ~ThreadGroup() {
    for (auto& t : threads_)
      if (t.joinable()) t.join();
  }

 private:
  std::vector<std::thread> threads_;
  std::mutex mutex_;
  std::exception_ptr error_;
};

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
struct TransposeLayout;
SellLayout sell_layout(const LowerPattern& lp, const TransposeLayout& t);  // t: lp's transpose

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

// One dataflow triangular solve in logical row order, built from the strictly-lower layout
// (off[n+1], col[]: a row's entries in the reference's summation order, ascending column),
// entry sources src into the factor's value array and diagonal sources dsrc.
//
// Rows are grouped into chains ("blocks") of at most max_rows rows that one wavefront solves one
// after the other, so that a dependency inside a chain costs no cross-wave hop. Rows are visited
// in increasing order; a row joins the chain of its critical predecessor (the dependency
// expected to finish last, by a model of CHAIN_HOP_US per cross-chain dependency and
// CHAIN_ROW_US per row) when that chain ends in it, has room (max_rows rows, xcap entries from
// outside the chain, icap from inside, ucap distinct outside columns), every other dependency
// of the row lies in an older chain (creation order is then topological) and the row's
// dependencies inside the chain are its last columns. Otherwise the row starts a new chain.
//
// A block's rows occupy consecutive "slots" (brow: the logical row of a slot, ascending within
// the block). Each slot's entries are split into the external part (columns in other blocks)
// and the internal part (the position, 0..63, of a row of the same block); since the internal
// columns are the row's last, externals then internals is the reference's order. Both parts
// are stored slot after slot, so a wave loads a whole block with coalesced reads.
//
// The rows of a chain are mesh neighbours and share most of their external columns: each block
// also lists its distinct external columns (ucol, ascending) and each external
// entry names its column by its index in that list (xu), so that a wave waits on, and loads,
// every external value once.
constexpr double CHAIN_HOP_US = 3.0;
constexpr double CHAIN_ROW_US = 0.25;

// Per-block capacities of a triangular-solve layout (the kernels stage each part in LDS):
// external entries, internal entries and distinct external columns (at most 256: xu is a byte).
struct TriCaps {
  int xcap, icap, ucap;
};

struct TriLayout {
  std::vector<int32_t> bstart;           // blocks + 1: slot offsets
  std::vector<int32_t> border;           // processing order: blocks sorted by block-DAG level
  std::vector<int32_t> brow;             // logical row of each slot
  std::vector<int32_t> xoff, xcol, xsrc;  // external entries: slots + 1, columns, sources
  std::vector<int32_t> ioff, ik, isrc;    // internal entries: slots + 1, positions, sources
  std::vector<int32_t> dsrc;              // diagonal source per slot
  std::vector<int32_t> uoff, ucol;        // distinct external columns: blocks + 1, columns
  std::vector<uint8_t> xu;                // per external entry: its column's index in ucol
  int max_block_ext = 0, max_block_int = 0, max_block_cols = 0;
};
TriLayout tri_layout(int64_t n, const int32_t* off, const int32_t* col, const int32_t* src,
                     const int32_t* dsrc, int max_rows, const TriCaps& caps, std::string* err);

// Every index array the device kernels of one operator read, built on the host from the lower
// pattern: the 32-bit CSR (IC(0)), the sliced-ELL symmetric layout (SpMV) and the two
// triangular-solve layouts. The backward solve L^T z = y runs in logical rows l = n-1-j: row l
// holds the entries l_ij, i > j, of column j of L in decreasing i, at logical column n-1-i.
struct DeviceLayouts {
  int64_t n = 0, nnz = 0, depth = 0, slices = 0;
  int max_row = 0;  // longest row of L, diagonal included
  std::vector<int32_t> off32, col32;
  // IC(0) processing order: the rows by level in the dependency DAG of L (a row's level is one
  // more than the highest level among its columns), ascending row within a level; it is
  // topological, and neighbouring positions are mostly independent rows
  std::vector<int32_t> ic_order;
  std::vector<int32_t> soff, scol, slen, ssrc;
  // col - row of every sliced-ELL entry as 16 bits when all fit (padding: 0), else empty
  std::vector<int16_t> sdel;
  // The same columns as one stream per group of rows: consecutive rows of a slice whose
  // column sets are equal (the DOFs of one node) share the ascending list of those columns,
  // as 16-bit offsets from the group's first row. Group g of slice sl with G groups and lists
  // of W entries holds entry q at goff[sl] + q G + g (W G entries per slice, padding 0). Per
  // row, ginfo packs the row length (bits 0-7), the position of the row's own column in its
  // group's list (8-15), the group's index in the slice (16-21) and the row's distance from
  // the group's first row (24-29). Row j's sliced-ELL entry k is list entry k - 1 below its
  // own position and k above it (k = 0 is the diagonal). Empty when some row is longer than
  // 255 entries or an offset does not fit 16 bits.
  std::vector<int32_t> goff;   // slices + 1
  std::vector<int32_t> gnum;   // groups per slice
  std::vector<int16_t> gdel;
  std::vector<uint32_t> ginfo;
  TriLayout fw, bw;
};
// Returns an empty string or the reason the pattern cannot be laid out (row too long).
std::string device_layouts(const LowerPattern& lp, const TriCaps& caps, DeviceLayouts* out);

// Verifies every bound the kernels rely on (index ranges, block capacities, coverage and order
// of each row's entries, topological block order); empty string when all hold.
std::string check_layouts(const LowerPattern& lp, const DeviceLayouts& d, const TriCaps& caps);

// Sliced ELL (64 rows per slice, entries column-major within a slice) of one strict triangle
// of the factor, for its Jacobi sweeps: row i holds the strictly lower entries l_ij, j < i, in
// ascending j (lower), or the strictly upper entries of L^T, l_ji for j > i in ascending j
// (upper). src is each entry's index in the lower-CSR value array; a padded position repeats
// the row's own column with src -1 and is never read (len bounds each row).
struct SweepLayout {
  int64_t slices = 0;
  std::vector<int32_t> soff;  // slices + 1, in entries
  std::vector<int32_t> scol, ssrc, slen, dsrc;  // dsrc[i]: the diagonal l_ii in lower-CSR order
  // |i - j| of every entry as 16 bits (0 for padding) when all fit, else empty: the kernel then
  // reads 2 bytes of column per entry instead of 4
  std::vector<uint16_t> sdelta;
};
// Both triangles' layouts; an empty string or the reason they do not fit 32-bit offsets.
std::string sweep_layouts(const LowerPattern& lp, SweepLayout* lower, SweepLayout* upper);
// Verifies every index the sweep kernel reads and that each row lists exactly its strict
// triangle, in ascending column; empty string when all hold.
std::string check_sweep_layouts(const LowerPattern& lp, const SweepLayout& lower,
                                const SweepLayout& upper);

// One strict triangle of the factor laid out for the fused sweep kernel, which runs every
// sweep of a block of `rows` (256 or 512) consecutive rows in one workgroup over a block of LDS
// that holds the block's own rows (slots 0..rows-1, row b*rows + s in slot s) and the distinct
// columns outside the block that its rows read (slots rows.., ascending column). Entries are
// stored by sub-block of FUSED_LANES rows (one row per lane of the workgroup) in quads: entry q
// of row 256 s + t at eoff[s] + (q >> 2) * 4 * FUSED_LANES + 4 t + (q & 3), so that a lane
// reads four entries with one load. Each entry has its coefficient's index in the lower-CSR
// value array (esrc; -1 for padding) and its LDS slot (eidx; padding names the row's own slot
// and is never summed). Rows list the same entries as SweepLayout, in ascending column: l_ij,
// j < i (lower) or l_ji, j > i (upper).
constexpr int FUSED_LANES = 256;
constexpr int FUSED_EXT = 1792;  // external columns a block may read
constexpr int FUSED_WIDTH = 252; // longest row (entries per row are counted in 8 bits)
struct FusedSweepLayout {
  int rows = 0;        // rows per block
  int64_t blocks = 0;  // blocks of `rows` rows
  int width = 0;       // the longest row
  std::vector<int32_t> eoff;  // sub-blocks + 1
  std::vector<int32_t> esrc;
  std::vector<uint16_t> eidx;
  std::vector<uint8_t> elen;  // per row
  std::vector<int32_t> xoff;  // blocks + 1
  std::vector<int32_t> xcol;
  int max_ext = 0;  // the most external columns of any block
};
// Both triangles' fused layouts in blocks of `rows` rows; an empty string, or the reason they
// do not fit (a row longer than FUSED_WIDTH, a block reading more than FUSED_EXT external
// columns, 32-bit offsets).
std::string fused_sweep_layouts(const LowerPattern& lp, int rows, FusedSweepLayout* lower,
                                FusedSweepLayout* upper);
// Verifies every index the fused kernel reads: each row's entries are its strict triangle in
// ascending column with the right sources, every slot names that column's LDS place, the
// external lists are ascending, distinct and outside the block, and every external column of
// a block lies in a block the kernel processes earlier (lower blocks for the lower triangle,
// higher ones for the upper); empty string when all hold.
std::string check_fused_sweep_layouts(const LowerPattern& lp, const FusedSweepLayout& lower,
                                      const FusedSweepLayout& upper);
// The fused kernel's sweeps restated on the host, block by block in its processing order and
// staging order: k sweeps of L y = r from y = r * rd, then k of L^T z = y from z = y * rd, with
// FP32 coefficients (l rounded to float), rd = 1 / (double)(float)l_ii and FP64 sums in each
// row's order. l holds the factor's values in lower-CSR order.
void fused_sweeps_host(const LowerPattern& lp, const FusedSweepLayout& lower,
                       const FusedSweepLayout& upper, const double* l, const double* r, int k,
                       double* z);

// The preconditioner z = (L L^T)^-1 r computed on the host through the triangular-solve
// layouts, block by block in processing order, each row summed in the order the device sums
// it. l holds the factor's values in lower-CSR order.
void layout_precondition(const DeviceLayouts& d, const double* l, const double* r, double* z);

}  // namespace ml
