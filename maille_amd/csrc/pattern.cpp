#include "pattern.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>

namespace ml {

namespace {

int worker_count() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 4;
  return (int)std::min(h, 16u);
}

// Runs body(begin, end) over [0, n) split into contiguous ranges, one per worker.
template <class F>
void parallel_ranges(i64 n, F body) {
  int t = worker_count();
  if (n < 4096 || t <= 1) {
    body((i64)0, n);
    return;
  }
  std::vector<std::thread> pool;
  i64 step = (n + t - 1) / t;
  for (int k = 0; k < t; ++k) {
    i64 b = k * step, e = std::min(n, b + step);
    if (b >= e) break;
    pool.emplace_back([=] { body(b, e); });
  }
  for (auto& th : pool) th.join();
}

}  // namespace

// iterative.rs:38-76. The neighbour order matters (RCM visits neighbours in list order), so
// it is the reference's: incident elements in element order, each element's vertices in list
// order, first occurrence kept, the vertex itself excluded. Rows are built in parallel; each
// worker keeps its own "already listed" marks.
Graph graph_from_elements(i64 n, i64 n_elements, const i64* eoff, const i64* eit) {
  std::vector<i64> incident(n + 1, 0);
  for (i64 e = 0; e < n_elements; ++e)
    for (i64 p = eoff[e]; p < eoff[e + 1]; ++p) incident[eit[p] + 1] += 1;
  for (i64 v = 0; v < n; ++v) incident[v + 1] += incident[v];
  std::vector<i64> cursor(incident.begin(), incident.end() - 1);
  std::vector<int32_t> of_vertex(incident[n]);
  for (i64 e = 0; e < n_elements; ++e)
    for (i64 p = eoff[e]; p < eoff[e + 1]; ++p) of_vertex[cursor[eit[p]]++] = (int32_t)e;

  Graph g;
  g.offsets.assign(n + 1, 0);
  parallel_ranges(n, [&](i64 b, i64 e) {
    std::vector<int32_t> mark(n, -1);
    for (i64 v = b; v < e; ++v) {
      mark[v] = (int32_t)v;
      i64 count = 0;
      for (i64 q = incident[v]; q < incident[v + 1]; ++q) {
        i64 el = of_vertex[q];
        for (i64 p = eoff[el]; p < eoff[el + 1]; ++p) {
          i64 w = eit[p];
          if (mark[w] != (int32_t)v) { mark[w] = (int32_t)v; ++count; }
        }
      }
      g.offsets[v + 1] = count;
    }
  });
  for (i64 v = 0; v < n; ++v) g.offsets[v + 1] += g.offsets[v];
  g.targets.resize(g.offsets[n]);
  parallel_ranges(n, [&](i64 b, i64 e) {
    std::vector<int32_t> mark(n, -1);
    for (i64 v = b; v < e; ++v) {
      mark[v] = (int32_t)v;
      i64 out = g.offsets[v];
      for (i64 q = incident[v]; q < incident[v + 1]; ++q) {
        i64 el = of_vertex[q];
        for (i64 p = eoff[el]; p < eoff[el + 1]; ++p) {
          i64 w = eit[p];
          if (mark[w] != (int32_t)v) { mark[w] = (int32_t)v; g.targets[out++] = (int32_t)w; }
        }
      }
    }
  });
  return g;
}

namespace {

// solve.rs:248-270 (bfs_levels) reduced to what the pseudo-peripheral search reads: the number
// of levels and the last level, in the reference's discovery order (a FIFO breadth-first
// search discovers each level in the same order as the level-by-level loop).
struct Eccentricity {
  i64 levels = 0;
  std::vector<i64> last;
};

Eccentricity bfs_eccentricity(const Graph& g, i64 start, const std::vector<char>& excluded,
                              std::vector<char>& seen, std::vector<i64>& queue) {
  std::memcpy(seen.data(), excluded.data(), excluded.size());
  seen[start] = 1;
  queue.clear();
  queue.push_back(start);
  i64 level_begin = 0, level_end = 1, levels = 1;
  for (;;) {
    for (i64 h = level_begin; h < level_end; ++h) {
      i64 v = queue[h];
      for (i64 p = g.offsets[v]; p < g.offsets[v + 1]; ++p) {
        i64 w = g.targets[p];
        if (!seen[w]) { seen[w] = 1; queue.push_back(w); }
      }
    }
    if ((i64)queue.size() == level_end) break;
    level_begin = level_end;
    level_end = (i64)queue.size();
    ++levels;
  }
  Eccentricity r;
  r.levels = levels;
  r.last.assign(queue.begin() + level_begin, queue.begin() + level_end);
  return r;
}

}  // namespace

// solve.rs:200-245. Per connected component: seed = first unvisited vertex of minimum degree;
// George-Liu pseudo-peripheral search (restart from the first minimum-degree vertex of the
// last level while the level count grows); breadth-first numbering with each vertex's
// unvisited neighbours stably sorted by degree; finally the whole order is reversed.
std::vector<i64> reverse_cuthill_mckee(const Graph& g) {
  const i64 n = g.n();
  std::vector<char> visited(n, 0), seen(n, 0);
  std::vector<i64> order, queue, next;
  order.reserve(n);
  queue.reserve(n);
  i64 scan_from = 0;  // every vertex below it is visited
  for (;;) {
    while (scan_from < n && visited[scan_from]) ++scan_from;
    i64 seed = -1;
    for (i64 v = scan_from; v < n; ++v)
      if (!visited[v] && (seed < 0 || g.degree(v) < g.degree(seed))) seed = v;
    if (seed < 0) break;
    i64 start = seed;
    Eccentricity ecc = bfs_eccentricity(g, start, visited, seen, queue);
    for (;;) {
      i64 far = -1;
      for (i64 v : ecc.last)
        if (far < 0 || g.degree(v) < g.degree(far)) far = v;
      if (far < 0) break;
      Eccentricity far_ecc = bfs_eccentricity(g, far, visited, seen, queue);
      if (far_ecc.levels <= ecc.levels) break;
      start = far;
      ecc = std::move(far_ecc);
    }
    i64 head = (i64)order.size();
    visited[start] = 1;
    order.push_back(start);
    while (head < (i64)order.size()) {
      i64 v = order[head++];
      next.clear();
      for (i64 p = g.offsets[v]; p < g.offsets[v + 1]; ++p)
        if (!visited[g.targets[p]]) next.push_back(g.targets[p]);
      std::stable_sort(next.begin(), next.end(),
                       [&](i64 a, i64 b) { return g.degree(a) < g.degree(b); });
      for (i64 w : next)
        if (!visited[w]) { visited[w] = 1; order.push_back(w); }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

i64 LowerPattern::find(i64 i, i64 j) const {
  if (j > i) std::swap(i, j);
  auto b = cols.begin() + offsets[i], e = cols.begin() + offsets[i + 1];
  auto it = std::lower_bound(b, e, (int32_t)j);
  if (it == e || *it != (int32_t)j) return -1;
  return it - cols.begin();
}

// iterative.rs:91-114 (Csr::zeros), rows built in parallel.
LowerPattern lower_pattern(const Graph& g, const std::vector<i64>& order,
                           const std::vector<i64>& rank) {
  const i64 n = (i64)order.size();
  LowerPattern lp;
  lp.offsets.assign(n + 1, 0);
  parallel_ranges(n, [&](i64 b, i64 e) {
    for (i64 r = b; r < e; ++r) {
      i64 v = order[r], c = 1;
      for (i64 p = g.offsets[v]; p < g.offsets[v + 1]; ++p) c += rank[g.targets[p]] < r;
      lp.offsets[r + 1] = c;
    }
  });
  for (i64 r = 0; r < n; ++r) lp.offsets[r + 1] += lp.offsets[r];
  lp.cols.resize(lp.offsets[n]);
  parallel_ranges(n, [&](i64 b, i64 e) {
    for (i64 r = b; r < e; ++r) {
      i64 v = order[r], out = lp.offsets[r];
      for (i64 p = g.offsets[v]; p < g.offsets[v + 1]; ++p) {
        i64 c = rank[g.targets[p]];
        if (c < r) lp.cols[out++] = (int32_t)c;
      }
      std::sort(lp.cols.begin() + lp.offsets[r], lp.cols.begin() + out);
      lp.cols[out] = (int32_t)r;
    }
  });
  return lp;
}

TransposeLayout transpose_layout(const LowerPattern& lp) {
  const i64 n = lp.n();
  TransposeLayout t;
  t.offsets.assign(n + 1, 0);
  for (i64 i = 0; i < n; ++i)
    for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p) t.offsets[lp.cols[p] + 1] += 1;
  for (i64 j = 0; j < n; ++j) t.offsets[j + 1] += t.offsets[j];
  std::vector<i64> cursor(t.offsets.begin(), t.offsets.end() - 1);
  t.rows.resize(t.offsets[n]);
  t.src.resize(t.offsets[n]);
  for (i64 i = 0; i < n; ++i)
    for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p) {
      i64 q = cursor[lp.cols[p]]++;
      t.rows[q] = (int32_t)i;
      t.src[q] = p;
    }
  return t;
}

SellLayout sell_layout(const LowerPattern& lp) {
  const i64 n = lp.n();
  TransposeLayout t = transpose_layout(lp);
  SellLayout s;
  s.n = n;
  s.slices = (n + SellLayout::C - 1) / SellLayout::C;
  s.slice_offsets.assign(s.slices + 1, 0);
  auto row_len = [&](i64 j) {
    return (lp.offsets[j + 1] - lp.offsets[j]) + (t.offsets[j + 1] - t.offsets[j]);
  };
  for (i64 sl = 0; sl < s.slices; ++sl) {
    i64 w = 0;
    for (i64 j = sl * SellLayout::C; j < std::min(n, (sl + 1) * SellLayout::C); ++j)
      w = std::max(w, row_len(j));
    s.slice_offsets[sl + 1] = s.slice_offsets[sl] + w * SellLayout::C;
  }
  s.cols.assign(s.slice_offsets[s.slices], 0);
  s.src.assign(s.slice_offsets[s.slices], -1);
  parallel_ranges(n, [&](i64 b, i64 e) {
    for (i64 j = b; j < e; ++j) {
      i64 sl = j / SellLayout::C, lane = j % SellLayout::C;
      i64 base = s.slice_offsets[sl] + lane, k = 0;
      auto put = [&](i64 col, i64 src) {
        s.cols[base + k * SellLayout::C] = (int32_t)col;
        s.src[base + k * SellLayout::C] = src;
        ++k;
      };
      i64 diag = lp.offsets[j + 1] - 1;
      put(j, diag);
      for (i64 p = lp.offsets[j]; p < diag; ++p) put(lp.cols[p], p);
      for (i64 q = t.offsets[j]; q < t.offsets[j + 1]; ++q) put(t.rows[q], t.src[q]);
      // padding keeps a valid column (the row itself) so no lane ever reads out of range
      i64 w = (s.slice_offsets[sl + 1] - s.slice_offsets[sl]) / SellLayout::C;
      for (; k < w;) {
        s.cols[base + k * SellLayout::C] = (int32_t)j;
        ++k;
      }
    }
  });
  return s;
}

i64 dag_depth(const LowerPattern& lp) {
  const i64 n = lp.n();
  std::vector<int32_t> level(n, 0);
  i64 depth = 0;
  for (i64 i = 0; i < n; ++i) {
    int32_t m = 0;
    for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p)
      m = std::max(m, level[lp.cols[p]] + 1);
    level[i] = m;
    depth = std::max<i64>(depth, m + 1);
  }
  return depth;
}

std::vector<int32_t> dataflow_blocks(int64_t n, const int32_t* off, const int32_t* col,
                                     int max_rows) {
  std::vector<int32_t> level(n, 0);
  std::vector<int32_t> starts;
  starts.reserve(n / 16 + 2);
  for (int64_t i = 0; i < n; ++i) {
    int32_t m = 0;
    for (int64_t p = off[i]; p < off[i + 1]; ++p) m = std::max(m, level[col[p]] + 1);
    level[i] = m;
    if (i == 0 || i - starts.back() >= max_rows || level[i] <= level[i - 1])
      starts.push_back((int32_t)i);
  }
  starts.push_back((int32_t)n);
  return starts;
}

TriLayout tri_layout(int64_t n, const int32_t* off, const int32_t* col, const int32_t* src,
                     const int32_t* dsrc, int max_rows, int xcap, int icap, std::string* err) {
  TriLayout t;
  std::vector<int32_t> level(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    int32_t m = 0;
    for (int64_t p = off[i]; p < off[i + 1]; ++p) m = std::max(m, level[col[p]] + 1);
    level[i] = m;
  }
  t.bstart.push_back(0);
  int64_t bx = 0, bi = 0;  // external / internal entries of the open block
  for (int64_t i = 0; i < n; ++i) {
    const int64_t len = off[i + 1] - off[i];
    if (len > xcap) {
      *err = "matrix row longer than the triangular-solve block capacity";
      return t;
    }
    if (i > 0) {
      const int64_t s = t.bstart.back();
      int64_t nin = 0;
      for (int64_t p = off[i]; p < off[i + 1]; ++p) nin += col[p] >= s;
      const bool cut = i - s >= max_rows || level[i] <= level[i - 1] ||
                       bx + (len - nin) > xcap || bi + nin > icap;
      if (cut) {
        t.bstart.push_back((int32_t)i);
        bx = 0;
        bi = 0;
      }
    }
    const int64_t s = t.bstart.back();
    for (int64_t p = off[i]; p < off[i + 1]; ++p) (col[p] >= s ? bi : bx) += 1;
    t.max_block_ext = std::max<int>(t.max_block_ext, (int)bx);
    t.max_block_int = std::max<int>(t.max_block_int, (int)bi);
  }
  t.bstart.push_back((int32_t)n);
  t.xoff.assign(n + 1, 0);
  t.ioff.assign(n + 1, 0);
  t.dsrc.assign(dsrc, dsrc + n);
  for (size_t b = 0; b + 1 < t.bstart.size(); ++b) {
    const int32_t s = t.bstart[b];
    for (int32_t i = s; i < t.bstart[b + 1]; ++i) {
      for (int64_t p = off[i]; p < off[i + 1]; ++p) {
        if (col[p] < s) {
          t.xcol.push_back(col[p]);
          t.xsrc.push_back(src[p]);
        } else {
          t.ik.push_back(col[p] - s);
          t.isrc.push_back(src[p]);
        }
      }
      t.xoff[i + 1] = (int32_t)t.xcol.size();
      t.ioff[i + 1] = (int32_t)t.ik.size();
    }
  }
  // Processing order: blocks by their level in the block dependency graph (a block depends
  // on the blocks holding its external columns), ties in index order. It is a topological
  // order, so a wave never waits for a block that comes after its own in the sequence.
  const int64_t nb = (int64_t)t.bstart.size() - 1;
  std::vector<int32_t> block_of(n), blev(nb, 0);
  for (int64_t b = 0; b < nb; ++b)
    for (int32_t i = t.bstart[b]; i < t.bstart[b + 1]; ++i) block_of[i] = (int32_t)b;
  for (int64_t b = 0; b < nb; ++b) {
    int32_t m = 0;
    for (int32_t q = t.xoff[t.bstart[b]]; q < t.xoff[t.bstart[b + 1]]; ++q)
      m = std::max(m, blev[block_of[t.xcol[q]]] + 1);
    blev[b] = m;
  }
  t.border.resize(nb);
  for (int64_t b = 0; b < nb; ++b) t.border[b] = (int32_t)b;
  std::stable_sort(t.border.begin(), t.border.end(),
                   [&](int32_t a, int32_t c) { return blev[a] < blev[c]; });
  return t;
}

}  // namespace ml
