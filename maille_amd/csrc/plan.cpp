#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>

namespace ml {

namespace {

const int TET10_EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

double distance3(const double* p, const double* q) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s += (p[k] - q[k]) * (p[k] - q[k]);
  return std::sqrt(s);
}

template <class F>
void parallel_for(i64 n, F body) {
  unsigned t = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (n < 65536 || t == 1) {
    body((i64)0, n);
    return;
  }
  ThreadGroup pool;
  i64 step = (n + t - 1) / t;
  for (unsigned k = 0; k < t; ++k) {
    i64 b = k * step, e = std::min(n, b + step);
    if (b >= e) break;
    pool.run([=] { body(b, e); });
  }
  pool.join();
}

// element_dofs (assembly.rs:20-32, Tet branch): [3n, 3n+1, 3n+2] per node, or the element's
// nodes when the model carries one DOF per node (temperature, model.rs:1074-1078)
inline i64 dof_of(const MeshView& m, i64 e, int local) {
  return m.dpn == 3 ? 3 * m.conn[e * m.nen + local / 3] + local % 3 : m.conn[e * m.nen + local];
}

}  // namespace

int check_mesh(const MeshView& m, i64* where) {
  *where = -1;
  if (m.n_nodes <= 0 || m.n_elements <= 0) return 7;
  for (i64 v = 0; v < m.n_nodes; ++v)
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(m.nodes[3 * v + k])) { *where = v; return 4; }
  for (i64 e = 0; e < m.n_elements; ++e) {
    // material.rs:145-153 (with_poisson) and Material::new's E > 0
    double E = m.young[e], nu = m.nu[e];
    if (!(std::isfinite(E) && E > 0.0) || !(nu > -1.0 && nu < 0.5)) { *where = e; return 6; }
  }
  for (i64 e = 0; e < m.n_elements; ++e) {
    const i64* t = m.conn + e * m.nen;
    for (int a = 0; a < m.nen; ++a)
      if (t[a] < 0 || t[a] >= m.n_nodes) { *where = t[a]; return 5; }  // UnknownNode { node }
    // mesh.rs:329-348 (check_tetrahedron)
    const double* c[10];
    for (int a = 0; a < m.nen; ++a) c[a] = m.nodes + 3 * t[a];
    double ev[3][3];
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j) ev[k][j] = c[k + 1][j] - c[0][j];
    double six_v = ev[0][0] * (ev[1][1] * ev[2][2] - ev[1][2] * ev[2][1]) -
                   ev[0][1] * (ev[1][0] * ev[2][2] - ev[1][2] * ev[2][0]) +
                   ev[0][2] * (ev[1][0] * ev[2][1] - ev[1][1] * ev[2][0]);
    double longest = 0.0;
    for (auto& ed : TET10_EDGES) longest = std::fmax(longest, distance3(c[ed[0]], c[ed[1]]));
    if (std::fabs(six_v) <= 1e-12 * (longest * longest * longest)) { *where = e; return 1; }
    if (six_v < 0.0) { *where = e; return 2; }
    // mesh.rs:401-419 (straight_edges) for Tet10
    if (m.nen == 10) {
      for (int k = 0; k < 6; ++k) {
        const double* a = c[TET10_EDGES[k][0]];
        const double* b = c[TET10_EDGES[k][1]];
        double mid[3] = {(a[0] + b[0]) / 2.0, (a[1] + b[1]) / 2.0, (a[2] + b[2]) / 2.0};
        if (!(distance3(c[4 + k], mid) <= 1e-9 * distance3(a, b))) { *where = e; return 3; }
      }
    }
  }
  return 0;
}

int build_plan(const MeshView& m, const uint8_t* fixed, const double* imposed, Plan& p,
               std::string* err) {
  auto t0 = std::chrono::steady_clock::now();
  p.n_nodes = m.n_nodes;
  p.n_elements = m.n_elements;
  p.nen = m.nen;
  p.dpn = m.dpn;
  p.m = m.dpn * m.nen;
  p.ndof = m.dpn * m.n_nodes;
  const int mm = p.m;
  const i64 ne = m.n_elements;
  if (p.ndof >= (i64)INT32_MAX || (i64)mm * mm * ne >= (i64)UINT32_MAX) {
    *err = "model too large for 32-bit device indices";
    return 3;
  }
  // model.rs:1449-1454: free DOFs numbered in increasing global order
  std::vector<i64> free_of(p.ndof, -1);
  i64 count = 0;
  for (i64 g = 0; g < p.ndof; ++g)
    if (!fixed[g]) free_of[g] = count++;
  p.nfree = count;  // may be 0: every DOF imposed, as FreeSystem::new allows
  // model.rs:1455-1458: per-element free DOFs in element_dofs order (assembly.rs:20-32)
  std::vector<i64> eoff(ne + 1, 0);
  for (i64 e = 0; e < ne; ++e) {
    i64 c = 0;
    for (int a = 0; a < mm; ++a) c += !fixed[dof_of(m, e, a)];
    eoff[e + 1] = eoff[e] + c;
  }
  std::vector<i64> eit(eoff[ne]);
  parallel_for(ne, [&](i64 b, i64 en) {
    for (i64 e = b; e < en; ++e) {
      i64 o = eoff[e];
      for (int a = 0; a < mm; ++a) {
        i64 f = free_of[dof_of(m, e, a)];
        if (f >= 0) eit[o++] = f;
      }
    }
  });
  // model.rs:1460-1465: Graph, RCM, rank
  Graph graph = graph_from_elements(count, ne, eoff.data(), eit.data());
  std::vector<i64>().swap(eit);
  p.order = reverse_cuthill_mckee(graph);
  std::vector<i64> rank(count);
  for (i64 r = 0; r < count; ++r) rank[p.order[r]] = r;
  p.lower = lower_pattern(graph, p.order, rank);
  Graph().offsets.swap(graph.offsets);
  std::vector<int32_t>().swap(graph.targets);
  if (p.lower.nnz() >= (i64)INT32_MAX) {
    *err = "too many matrix entries for 32-bit device indices";
    return 3;
  }
  p.depth = dag_depth(p.lower);
  p.rank_of_dof.assign(p.ndof, -1);
  for (i64 g = 0; g < p.ndof; ++g)
    if (free_of[g] >= 0) p.rank_of_dof[g] = (int32_t)rank[free_of[g]];

  // model.rs:1480-1492 scatter order: element e, row li, column lj; the entry (rank[a],
  // rank[b]) with rank[b] <= rank[a] receives ke[li][lj] (the row is the DOF of larger rank).
  // Each element reaches an entry at most once, so an entry's contributions listed in
  // ascending element order keep the reference's order. They are built row by row, rows in
  // parallel: row ra walks the elements of its DOF's node in ascending order, and for each the
  // element's DOFs of rank <= ra (two passes: counts per entry, then slots), so that each entry
  // is written only by the thread that owns its row.
  {
    const int nen = m.nen, dpn = m.dpn;
    std::vector<int64_t> nel_off(m.n_nodes + 1, 0);
    for (i64 e = 0; e < ne; ++e)
      for (int a = 0; a < nen; ++a) nel_off[m.conn[e * nen + a] + 1] += 1;
    for (i64 v = 0; v < m.n_nodes; ++v) nel_off[v + 1] += nel_off[v];
    std::vector<int64_t> nel_item(nel_off[m.n_nodes]);  // e * nen + position of the node
    {
      std::vector<int64_t> cursor(nel_off.begin(), nel_off.end() - 1);
      for (i64 e = 0; e < ne; ++e)
        for (int a = 0; a < nen; ++a) nel_item[cursor[m.conn[e * nen + a]]++] = e * nen + a;
    }
    std::vector<i64> dof_of_rank(count);
    for (i64 g = 0; g < p.ndof; ++g)
      if (p.rank_of_dof[g] >= 0) dof_of_rank[p.rank_of_dof[g]] = g;
    // visits (entry q, Ke slot) of row ra in ascending element order; false on an entry
    // outside the pattern
    auto row_entries = [&](i64 ra, auto&& visit) {
      const i64 g = dof_of_rank[ra];
      const i64 v = g / dpn;
      const int c = (int)(g % dpn);
      for (int64_t t = nel_off[v]; t < nel_off[v + 1]; ++t) {
        const i64 e = nel_item[t] / nen;
        const int la = dpn * (int)(nel_item[t] % nen) + c;
        for (int lj = 0; lj < mm; ++lj) {
          const int32_t rb = p.rank_of_dof[dof_of(m, e, lj)];
          if (rb < 0 || rb > ra) continue;
          const i64 q = p.lower.find(ra, rb);
          if (q < 0) return false;
          visit(q, (uint32_t)((i64)(la * mm + lj) * ne + e));
        }
      }
      return true;
    };
    std::vector<int64_t> counts(p.lower.nnz() + 1, 0);
    std::atomic<bool> bad{false};
    parallel_for(count, [&](i64 b, i64 en) {
      for (i64 ra = b; ra < en; ++ra)
        if (!row_entries(ra, [&](i64 q, uint32_t) { counts[q + 1] += 1; })) bad = true;
    });
    if (bad) {
      *err = "element entry outside the pattern";
      return 4;
    }
    for (i64 q = 0; q < p.lower.nnz(); ++q) counts[q + 1] += counts[q];
    p.asm_slot.resize(counts[p.lower.nnz()]);
    parallel_for(count, [&](i64 b, i64 en) {
      std::vector<int64_t> cursor;
      for (i64 ra = b; ra < en; ++ra) {
        const i64 q0 = p.lower.offsets[ra];
        cursor.assign(counts.begin() + q0, counts.begin() + p.lower.offsets[ra + 1]);
        row_entries(ra, [&](i64 q, uint32_t slot) { p.asm_slot[cursor[q - q0]++] = slot; });
      }
    });
    p.asm_off.swap(counts);
  }

  // coupling K_fc u_c (model.rs:1489): row a free, column gj fixed, element then lj order
  bool any_imposed = false;
  for (i64 g = 0; g < p.ndof; ++g) any_imposed = any_imposed || (fixed[g] && imposed[g] != 0.0);
  p.cpl_off.assign(count + 1, 0);
  if (any_imposed) {
    for (i64 e = 0; e < ne; ++e)
      for (int li = 0; li < mm; ++li) {
        int32_t ra = p.rank_of_dof[dof_of(m, e, li)];
        if (ra < 0) continue;
        for (int lj = 0; lj < mm; ++lj) {
          i64 gj = dof_of(m, e, lj);
          if (fixed[gj]) p.cpl_off[ra + 1] += 1;
        }
      }
    for (i64 a = 0; a < count; ++a) p.cpl_off[a + 1] += p.cpl_off[a];
    p.cpl_slot.resize(p.cpl_off[count]);
    p.cpl_dof.resize(p.cpl_off[count]);
    std::vector<int64_t> cursor(p.cpl_off.begin(), p.cpl_off.end() - 1);
    for (i64 e = 0; e < ne; ++e)
      for (int li = 0; li < mm; ++li) {
        int32_t ra = p.rank_of_dof[dof_of(m, e, li)];
        if (ra < 0) continue;
        for (int lj = 0; lj < mm; ++lj) {
          i64 gj = dof_of(m, e, lj);
          if (!fixed[gj]) continue;
          i64 c = cursor[ra]++;
          p.cpl_slot[c] = (uint32_t)((i64)(li * mm + lj) * ne + e);
          p.cpl_dof[c] = (int32_t)gj;
        }
      }
  }

  // reactions (model.rs:1353-1369): fixed DOFs in global order, rows (e, li) in element order
  std::vector<int32_t> fixed_index(p.ndof, -1);
  for (i64 g = 0; g < p.ndof; ++g)
    if (fixed[g]) {
      fixed_index[g] = (int32_t)p.rea_dof.size();
      p.rea_dof.push_back((int32_t)g);
    }
  p.rea_off.assign(p.rea_dof.size() + 1, 0);
  for (i64 e = 0; e < ne; ++e)
    for (int li = 0; li < mm; ++li) {
      int32_t k = fixed_index[dof_of(m, e, li)];
      if (k >= 0) p.rea_off[k + 1] += 1;
    }
  for (size_t k = 0; k < p.rea_dof.size(); ++k) p.rea_off[k + 1] += p.rea_off[k];
  p.rea_item.resize(p.rea_off[p.rea_dof.size()]);
  {
    std::vector<int64_t> cursor(p.rea_off.begin(), p.rea_off.end() - 1);
    for (i64 e = 0; e < ne; ++e)
      for (int li = 0; li < mm; ++li) {
        int32_t k = fixed_index[dof_of(m, e, li)];
        if (k >= 0) p.rea_item[cursor[k]++] = e * mm + li;
      }
  }
  p.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

}  // namespace ml
