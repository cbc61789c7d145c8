// Host plan of one static solve (model.rs:1070-1147 and 1442-1499): mesh checks, free-DOF
// numbering, the reverse Cuthill-McKee lower pattern and every index list the device needs to
// assemble K_ff, the coupling K_fc u_c and the reactions in the reference's summation order.
#pragma once

#include <cstdint>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "pattern.hpp"

namespace ml {

struct MeshView {
  i64 n_nodes = 0, n_elements = 0;
  int nen = 0;                    // 4 (Tet4) or 10 (Tet10)
  const double* nodes = nullptr;  // n_nodes x 3
  const i64* conn = nullptr;      // n_elements x nen
  const double* young = nullptr;
  const double* nu = nullptr;
  int dpn = 3;                    // DOFs per node: 3 (displacements) or 1 (temperature)
};

// mesh.rs:257-348 + material.rs:145-153: returns 0 or a mesh error code (see maille_hip.h)
// with the offending element (or node) in *where.
int check_mesh(const MeshView& m, i64* where);

// std::allocator whose value-less construct() leaves the element uninitialised, so that
// resizing a large index array that is written in full next costs no serial zero fill.
template <class T>
struct DefaultInitAllocator : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = DefaultInitAllocator<U>;
  };
  DefaultInitAllocator() = default;
  template <class U>
  DefaultInitAllocator(const DefaultInitAllocator<U>&) noexcept {}
  template <class U>
  void construct(U* p) {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};

struct Plan {
  i64 n_nodes = 0, n_elements = 0, ndof = 0, nfree = 0;
  int nen = 0, m = 0;  // m = dpn nen
  int dpn = 3;
  std::vector<int32_t> rank_of_dof;  // global DOF -> solver row (RCM rank) or -1 if fixed
  std::vector<i64> order;            // RCM order over free indices (order[new] = old)
  LowerPattern lower;
  i64 depth = 0;                     // DAG depth of the triangular solves
  // K_ff entry q = sum over asm_slot[asm_off[q] .. asm_off[q+1]) of Ke slots, element order.
  // A Ke slot is (j*m + i) * n_elements + e for ke[e][j][i] (entry-major, element fastest).
  std::vector<int64_t> asm_off;
  std::vector<uint32_t, DefaultInitAllocator<uint32_t>> asm_slot;  // every element written
  // coupling (K_fc u_c) of solver row a: sum of ke[slot] * imposed[dof] over
  // cpl_off[a]..cpl_off[a+1], element order; empty when every imposed value is zero.
  std::vector<int64_t> cpl_off;
  std::vector<uint32_t> cpl_slot;
  std::vector<int32_t> cpl_dof;
  // reactions: fixed DOF k (in increasing global order, rea_dof[k]) sums, in element order,
  // the rows (e, li) listed in rea_item[rea_off[k] .. rea_off[k+1]) as e * m + li.
  std::vector<int32_t> rea_dof;
  std::vector<int64_t> rea_off;
  std::vector<int64_t> rea_item;
  double seconds = 0.0;
};

// FreeSystem::new up to the scatter (model.rs:1442-1499), on the host. fixed: dpn n_nodes flags.
// Returns 0 or an ml_status code with a message in *err.
int build_plan(const MeshView& m, const uint8_t* fixed, const double* imposed, Plan& plan,
               std::string* err);

}  // namespace ml
