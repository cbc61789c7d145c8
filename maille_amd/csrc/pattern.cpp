#include "pattern.hpp"

#include <cstdlib>

#include <algorithm>
#include <atomic>
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
  ThreadGroup pool;
  i64 step = (n + t - 1) / t;
  for (int k = 0; k < t; ++k) {
    i64 b = k * step, e = std::min(n, b + step);
    if (b >= e) break;
    pool.run([=] { body(b, e); });
  }
  pool.join();
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

SellLayout sell_layout(const LowerPattern& lp) { return sell_layout(lp, transpose_layout(lp)); }

SellLayout sell_layout(const LowerPattern& lp, const TransposeLayout& t) {
  const i64 n = lp.n();
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

TriLayout tri_layout(int64_t n, const int32_t* off, const int32_t* col, const int32_t* src,
                     const int32_t* dsrc, int max_rows, const TriCaps& caps, std::string* err) {
  TriLayout t;
  const int xcap = caps.xcap, icap = caps.icap, ucap = caps.ucap;
  if (ucap < 1 || ucap > 256) {
    *err = "distinct-column capacity outside 1..256";
    return t;
  }
  // chains: rows in increasing order; a row joins the chain of its critical predecessor when
  // that chain ends in it, has room (rows, external and internal entries), every other
  // dependency lies in an older chain, and the row's entries inside the chain are the last of
  // its columns (so a row reads as external entries, then internal ones, in its own order)
  std::vector<int32_t> chain_of(n, -1), clen, clast, cext, cint;
  std::vector<std::vector<int32_t>> cdist;  // per chain: its distinct outside columns, ascending
  std::vector<int32_t> merged;
  std::vector<double> finish(n, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t len = off[i + 1] - off[i];
    if (len > xcap || len > icap || len > ucap) {
      *err = "matrix row longer than the triangular-solve block capacity";
      return t;
    }
    int32_t best = -1, newest = -1;
    double bf = -1.0;
    for (int64_t p = off[i]; p < off[i + 1]; ++p) {
      const int32_t j = col[p];
      if (finish[j] > bf) { bf = finish[j]; best = j; }
      newest = std::max(newest, chain_of[j]);
    }
    int32_t ch = -1;
    if (best >= 0) {
      const int32_t cb = chain_of[best];
      if (clast[cb] == best && clen[cb] < max_rows && newest <= cb) {
        int64_t ext = 0, in = 0;
        bool suffix = true;
        for (int64_t p = off[i]; p < off[i + 1]; ++p) {
          if (chain_of[col[p]] == cb) {
            ++in;
          } else {
            ++ext;
            if (in > 0) suffix = false;
          }
        }
        if (suffix && cext[cb] + ext <= xcap && cint[cb] + in <= icap) {
          // a row's columns are distinct, so its new outside columns are those not listed yet
          const std::vector<int32_t>& dl = cdist[cb];
          int64_t fresh = 0;
          for (int64_t p = off[i]; p < off[i + 1]; ++p)
            if (chain_of[col[p]] != cb && !std::binary_search(dl.begin(), dl.end(), col[p]))
              ++fresh;
          if ((int64_t)dl.size() + fresh <= ucap) ch = cb;
        }
      }
    }
    if (ch < 0) {
      ch = (int32_t)clen.size();
      clen.push_back(0);
      clast.push_back(-1);
      cext.push_back(0);
      cint.push_back(0);
      cdist.emplace_back();
    }
    double f = clast[ch] >= 0 ? finish[clast[ch]] + CHAIN_ROW_US : CHAIN_ROW_US;
    merged.clear();
    for (int64_t p = off[i]; p < off[i + 1]; ++p) {
      const int32_t j = col[p];
      if (chain_of[j] != ch) {
        f = std::max(f, finish[j] + CHAIN_HOP_US + CHAIN_ROW_US);
        cext[ch] += 1;
        merged.push_back(j);  // ascending: the row's columns are
      } else {
        cint[ch] += 1;
      }
    }
    if (!merged.empty()) {
      std::vector<int32_t>& dl = cdist[ch];
      const size_t old_size = dl.size();
      dl.insert(dl.end(), merged.begin(), merged.end());
      std::inplace_merge(dl.begin(), dl.begin() + old_size, dl.end());
      dl.erase(std::unique(dl.begin(), dl.end()), dl.end());
    }
    finish[i] = f;
    chain_of[i] = ch;
    clen[ch] += 1;
    clast[ch] = (int32_t)i;
  }
  // blocks = chains in creation order; their rows become consecutive slots
  const int64_t nb = (int64_t)clen.size();
  t.bstart.assign(nb + 1, 0);
  for (int64_t b = 0; b < nb; ++b) t.bstart[b + 1] = t.bstart[b] + clen[b];
  t.brow.resize(n);
  std::vector<int32_t> pos(n), fill(t.bstart.begin(), t.bstart.end() - 1);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t q = fill[chain_of[i]]++;
    t.brow[q] = (int32_t)i;
    pos[i] = q - t.bstart[chain_of[i]];
  }
  t.xoff.assign(n + 1, 0);
  t.ioff.assign(n + 1, 0);
  t.dsrc.resize(n);
  for (int64_t q = 0; q < n; ++q) {
    const int32_t i = t.brow[q];
    for (int64_t p = off[i]; p < off[i + 1]; ++p) {
      if (chain_of[col[p]] != chain_of[i]) {
        t.xcol.push_back(col[p]);
        t.xsrc.push_back(src[p]);
      } else {
        t.ik.push_back(pos[col[p]]);
        t.isrc.push_back(src[p]);
      }
    }
    t.xoff[q + 1] = (int32_t)t.xcol.size();
    t.ioff[q + 1] = (int32_t)t.ik.size();
    t.dsrc[q] = dsrc[i];
  }
  cdist.clear();
  cdist.shrink_to_fit();
  // each block's distinct outside columns, ascending (neighbouring lanes then load from the
  // same cache lines, which a wave's load merges), and every external entry's index in that
  // list
  t.uoff.assign(nb + 1, 0);
  t.xu.resize(t.xcol.size());
  std::vector<int32_t> uidx(n, -1);
  for (int64_t b = 0; b < nb; ++b) {
    const int32_t ub = (int32_t)t.ucol.size();
    for (int32_t x = t.xoff[t.bstart[b]]; x < t.xoff[t.bstart[b + 1]]; ++x) {
      const int32_t c = t.xcol[x];
      if (uidx[c] < 0) {
        uidx[c] = 0;
        t.ucol.push_back(c);
      }
    }
    std::sort(t.ucol.begin() + ub, t.ucol.end());
    for (int32_t u = ub; u < (int32_t)t.ucol.size(); ++u) uidx[t.ucol[u]] = u - ub;
    for (int32_t x = t.xoff[t.bstart[b]]; x < t.xoff[t.bstart[b + 1]]; ++x)
      t.xu[x] = (uint8_t)uidx[t.xcol[x]];
    for (int32_t u = ub; u < (int32_t)t.ucol.size(); ++u) uidx[t.ucol[u]] = -1;
    t.uoff[b + 1] = (int32_t)t.ucol.size();
  }
  for (int64_t b = 0; b < nb; ++b) {
    t.max_block_ext = std::max(t.max_block_ext, t.xoff[t.bstart[b + 1]] - t.xoff[t.bstart[b]]);
    t.max_block_int = std::max(t.max_block_int, t.ioff[t.bstart[b + 1]] - t.ioff[t.bstart[b]]);
    t.max_block_cols = std::max(t.max_block_cols, t.uoff[b + 1] - t.uoff[b]);
  }
  // Processing order: blocks by their level in the block dependency graph (a block depends
  // on the blocks holding its external columns), ties in creation order. It is a topological
  // order, so a wave never waits for a block that comes after its own in the sequence.
  std::vector<int32_t> blev(nb, 0);
  for (int64_t b = 0; b < nb; ++b) {
    int32_t m = 0;
    for (int32_t q = t.xoff[t.bstart[b]]; q < t.xoff[t.bstart[b + 1]]; ++q)
      m = std::max(m, blev[chain_of[t.xcol[q]]] + 1);
    blev[b] = m;
  }
  t.border.resize(nb);
  for (int64_t b = 0; b < nb; ++b) t.border[b] = (int32_t)b;
  std::stable_sort(t.border.begin(), t.border.end(),
                   [&](int32_t a, int32_t c) { return blev[a] < blev[c]; });
  return t;
}

// The grouped column streams of the sliced ELL (DeviceLayouts::goff ...), built from its
// per-row columns; left empty when some row or offset does not fit their fields. Slices are
// built in contiguous ranges, one per worker, each into a buffer of its own; the buffers are
// then joined in slice order.
static void grouped_columns(DeviceLayouts& d) {
  const i64 n = d.n;
  d.goff.clear(); d.gnum.clear(); d.gdel.clear(); d.ginfo.clear();
  std::vector<int32_t> gnum(d.slices, 0);
  std::vector<uint32_t> ginfo(n, 0);
  struct Part {
    std::vector<int16_t> gdel;
    std::vector<i64> ends;  // the buffer's length after each of the range's slices
    bool ok = true;
  };
  auto build = [&](i64 s0, i64 s1, Part& part) {
    std::vector<int32_t> next;
    auto row_cols = [&](i64 j, std::vector<int32_t>& out) {  // the row's columns, ascending
      const i64 base = d.soff[j >> 6] + (j & 63);
      out.resize(d.slen[j]);
      for (i64 k = 0; k < d.slen[j]; ++k) out[k] = d.scol[base + (k << 6)];
      std::sort(out.begin(), out.end());
    };
    std::vector<i64> first;  // first row of each group
    std::vector<std::vector<int32_t>> lists;
    for (i64 sl = s0; sl < s1; ++sl) {
      const i64 j0 = sl * 64, j1 = std::min<i64>(n, j0 + 64);
      first.clear();
      lists.clear();
      for (i64 j = j0; j < j1; ++j) {
        if (d.slen[j] > 255) { part.ok = false; return; }
        row_cols(j, next);
        if (!first.empty() && next == lists.back() && j - first.back() < 64) {
          // same column set as the group in progress
        } else {
          first.push_back(j);
          lists.push_back(next);
        }
        const i64 g = (i64)first.size() - 1;
        const auto& L = lists.back();
        const i64 pos = std::lower_bound(L.begin(), L.end(), (int32_t)j) - L.begin();
        if (pos >= (i64)L.size() || L[pos] != j || pos > 255) { part.ok = false; return; }
        ginfo[j] = (uint32_t)d.slen[j] | ((uint32_t)pos << 8) | ((uint32_t)g << 16) |
                   ((uint32_t)(j - first.back()) << 24);
      }
      const i64 G = (i64)first.size();
      i64 W = 0;
      for (const auto& L : lists) W = std::max<i64>(W, (i64)L.size());
      gnum[sl] = (int32_t)G;
      const i64 at = (i64)part.gdel.size();
      if (at + W * G > INT32_MAX) { part.ok = false; return; }
      part.gdel.resize(at + W * G, 0);
      for (i64 g = 0; g < G; ++g)
        for (i64 q = 0; q < (i64)lists[g].size(); ++q) {
          const i64 off = (i64)lists[g][q] - first[g];
          if (off < -32768 || off > 32767) { part.ok = false; return; }
          part.gdel[at + q * G + g] = (int16_t)off;
        }
      part.ends.push_back((i64)part.gdel.size());
    }
  };
  const i64 T = d.slices < 256 ? 1 : worker_count();
  const i64 step = (d.slices + T - 1) / std::max<i64>(T, 1);
  std::vector<Part> parts(T);
  {
    ThreadGroup pool;
    for (i64 k = 0; k < T; ++k) {
      const i64 s0 = std::min(d.slices, k * step), s1 = std::min(d.slices, s0 + step);
      pool.run([&, s0, s1, k] { build(s0, s1, parts[k]); });
    }
    pool.join();
  }
  std::vector<int32_t> goff(d.slices + 1, 0);
  i64 total = 0;
  for (i64 k = 0; k < T; ++k) {
    if (!parts[k].ok) return;
    const i64 s0 = std::min(d.slices, k * step);
    for (size_t u = 0; u < parts[k].ends.size(); ++u) {
      if (total + parts[k].ends[u] > INT32_MAX) return;
      goff[s0 + (i64)u + 1] = (int32_t)(total + parts[k].ends[u]);
    }
    total += (i64)parts[k].gdel.size();
  }
  std::vector<int16_t> gdel(total);
  for (i64 k = 0, at = 0; k < T; ++k) {
    std::copy(parts[k].gdel.begin(), parts[k].gdel.end(), gdel.begin() + at);
    at += (i64)parts[k].gdel.size();
  }
  d.goff.swap(goff);
  d.gnum.swap(gnum);
  d.gdel.swap(gdel);
  d.ginfo.swap(ginfo);
}

std::string device_layouts(const LowerPattern& lp, const TriCaps& caps, DeviceLayouts* out) {
  DeviceLayouts& d = *out;
  const i64 n = lp.n();
  d.n = n;
  d.nnz = lp.nnz();
  d.depth = dag_depth(lp);
  d.max_row = 0;
  for (i64 i = 0; i < n; ++i)
    d.max_row = std::max<int>(d.max_row, (int)(lp.offsets[i + 1] - lp.offsets[i]));
  d.off32.assign(lp.offsets.begin(), lp.offsets.end());
  d.col32.assign(lp.cols.begin(), lp.cols.end());
  {
    // rows by DAG level (counting sort: stable, so ascending row within a level)
    std::vector<int32_t> level(n, 0);
    int32_t top = 0;
    for (i64 i = 0; i < n; ++i) {
      int32_t m = 0;
      for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p) m = std::max(m, level[lp.cols[p]] + 1);
      level[i] = m;
      top = std::max(top, m);
    }
    std::vector<int64_t> at(top + 2, 0);
    for (i64 i = 0; i < n; ++i) at[level[i] + 1] += 1;
    for (int32_t v = 0; v <= top; ++v) at[v + 1] += at[v];
    d.ic_order.resize(n);
    for (i64 i = 0; i < n; ++i) d.ic_order[at[level[i]]++] = (int32_t)i;
  }
  // sliced ELL of the full symmetric matrix
  TransposeLayout t = transpose_layout(lp);
  {
    SellLayout s = sell_layout(lp, t);
    d.slices = s.slices;
    d.soff.assign(s.slice_offsets.begin(), s.slice_offsets.end());
    d.scol.swap(s.cols);
    d.ssrc.resize(s.src.size());
    parallel_ranges((i64)s.src.size(), [&](i64 b, i64 e) {
      for (i64 q = b; q < e; ++q) d.ssrc[q] = (int32_t)s.src[q];
    });
  }
  d.slen.resize(n);
  for (i64 j = 0; j < n; ++j)
    d.slen[j] = (int32_t)((lp.offsets[j + 1] - lp.offsets[j]) + (t.offsets[j + 1] - t.offsets[j]));
  {
    // 16-bit column offsets from the row, when the matrix band allows (slices in parallel)
    std::atomic<bool> fits{true};
    std::vector<int16_t> del(d.scol.size(), 0);
    parallel_ranges(d.slices, [&](i64 b, i64 e) {
      for (i64 sl = b; sl < e && fits; ++sl) {
        const i64 w = ((i64)d.soff[sl + 1] - d.soff[sl]) / SellLayout::C;
        for (i64 k = 0; k < w; ++k)
          for (i64 lane = 0; lane < SellLayout::C; ++lane) {
            const i64 j = sl * SellLayout::C + lane;
            const i64 at = d.soff[sl] + k * SellLayout::C + lane;
            const i64 dlt = j < n ? (i64)d.scol[at] - j : 0;
            if (dlt < -32768 || dlt > 32767) fits = false;
            del[at] = (int16_t)dlt;
          }
      }
    });
    if (fits) d.sdel.swap(del);
  }
  grouped_columns(d);
  // forward solve: the strictly lower rows of L
  const i64 nf = lp.nnz() - n;
  std::vector<int32_t> foff(n + 1), fcol(nf), fsrc(nf), fdsrc(n);
  for (i64 i = 0; i <= n; ++i) foff[i] = (int32_t)(lp.offsets[i] - i);
  parallel_ranges(n, [&](i64 b, i64 e) {
    for (i64 i = b; i < e; ++i) {
      i64 o = foff[i];
      for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p, ++o) {
        fcol[o] = lp.cols[p];
        fsrc[o] = (int32_t)p;
      }
      fdsrc[i] = (int32_t)(lp.offsets[i + 1] - 1);
    }
  });
  // backward solve in logical rows l = n-1-j
  std::vector<int32_t> boff(n + 1, 0), bcol(nf), bsrc(nf), bdsrc(n);
  for (i64 l = 0; l < n; ++l) {
    const i64 j = n - 1 - l;
    boff[l + 1] = boff[l] + (int32_t)(t.offsets[j + 1] - t.offsets[j]);
  }
  parallel_ranges(n, [&](i64 b, i64 e) {
    for (i64 l = b; l < e; ++l) {
      const i64 j = n - 1 - l;
      i64 o = boff[l];
      for (i64 q = t.offsets[j + 1] - 1; q >= t.offsets[j]; --q, ++o) {
        bcol[o] = (int32_t)(n - 1 - t.rows[q]);
        bsrc[o] = (int32_t)t.src[q];
      }
      bdsrc[l] = (int32_t)(lp.offsets[j + 1] - 1);
    }
  });
  TransposeLayout().offsets.swap(t.offsets);
  std::vector<int32_t>().swap(t.rows);
  std::vector<int64_t>().swap(t.src);
  // the two solves' layouts are independent: the backward one on a thread of its own
  std::string err, err_bw;
  ThreadGroup bw_thread;
  bw_thread.run([&] {
    d.bw = tri_layout(n, boff.data(), bcol.data(), bsrc.data(), bdsrc.data(), 64, caps, &err_bw);
  });
  d.fw = tri_layout(n, foff.data(), fcol.data(), fsrc.data(), fdsrc.data(), 64, caps, &err);
  bw_thread.join();
  return err.empty() ? err_bw : err;
}

namespace {

// One triangular-solve layout against the strictly-lower logical rows it was built from.
std::string check_tri(const char* name, const TriLayout& t, i64 n, i64 nnz, const TriCaps& caps,
                      const std::vector<int32_t>& rows_off, const std::vector<int32_t>& rows_col,
                      const std::vector<int32_t>& rows_src) {
  const std::string pre = std::string(name) + " layout: ";
  if (t.bstart.size() < 2 || t.bstart.front() != 0 || t.bstart.back() != n)
    return pre + "blocks do not span the rows";
  const i64 nb = (i64)t.bstart.size() - 1;
  if ((i64)t.brow.size() != n || (i64)t.xoff.size() != n + 1 || (i64)t.ioff.size() != n + 1 ||
      (i64)t.dsrc.size() != n)
    return pre + "slot arrays of the wrong length";
  if (t.xoff[0] != 0 || t.ioff[0] != 0 || (size_t)t.xoff[n] != t.xcol.size() ||
      t.xcol.size() != t.xsrc.size() || (size_t)t.ioff[n] != t.ik.size() ||
      t.ik.size() != t.isrc.size())
    return pre + "entry arrays do not match their offsets";
  if ((i64)t.uoff.size() != nb + 1 || t.uoff[0] != 0 || (size_t)t.uoff[nb] != t.ucol.size() ||
      t.xu.size() != t.xcol.size())
    return pre + "distinct-column arrays do not match their offsets";
  if (caps.ucap > 256) return pre + "distinct-column capacity over 256";
  std::vector<int32_t> block_of(n, -1), pos(nb, -1);
  for (i64 b = 0; b < nb; ++b) {
    const i64 rows = (i64)t.bstart[b + 1] - t.bstart[b];
    if (rows < 1 || rows > 64) return pre + "block of " + std::to_string(rows) + " rows";
    for (i64 q = t.bstart[b]; q < t.bstart[b + 1]; ++q) {
      const i64 i = t.brow[q];
      if (i < 0 || i >= n || block_of[i] >= 0) return pre + "slots are not a permutation of the rows";
      if (q > t.bstart[b] && t.brow[q] <= t.brow[q - 1]) return pre + "block rows not ascending";
      block_of[i] = (int32_t)b;
    }
  }
  if ((i64)t.border.size() != nb) return pre + "processing order of the wrong length";
  for (i64 k = 0; k < nb; ++k) {
    const i64 b = t.border[k];
    if (b < 0 || b >= nb || pos[b] >= 0) return pre + "processing order is not a permutation";
    pos[b] = (int32_t)k;
  }
  std::vector<char> listed(n, 0);
  for (i64 b = 0; b < nb; ++b) {
    const i64 s = t.bstart[b], e = t.bstart[b + 1];
    if (t.xoff[e] - t.xoff[s] > caps.xcap) return pre + "block over the external capacity";
    if (t.ioff[e] - t.ioff[s] > caps.icap) return pre + "block over the internal capacity";
    const i64 ub = t.uoff[b], nu = (i64)t.uoff[b + 1] - ub;
    if (nu < 0 || nu > caps.ucap) return pre + "block over the distinct-column capacity";
    for (i64 u = ub; u < ub + nu; ++u) {
      const i64 c = t.ucol[u];
      if (c < 0 || c >= n || listed[c]) return pre + "distinct columns out of range or repeated";
      listed[c] = 1;
    }
    for (i64 u = ub; u < ub + nu; ++u) listed[t.ucol[u]] = 0;
    for (i64 q = s; q < e; ++q) {
      const i64 i = t.brow[q], k = q - s;
      if (t.xoff[q + 1] < t.xoff[q] || t.ioff[q + 1] < t.ioff[q])
        return pre + "decreasing offsets at slot " + std::to_string(q);
      if (t.dsrc[q] < 0 || t.dsrc[q] >= nnz) return pre + "diagonal source out of range";
      // the slot's entries, externals then internals, are its row's logical entries in order
      i64 p = rows_off[i];
      if ((t.xoff[q + 1] - t.xoff[q]) + (t.ioff[q + 1] - t.ioff[q]) != rows_off[i + 1] - p)
        return pre + "row " + std::to_string(i) + " does not hold its entries";
      for (i64 x = t.xoff[q]; x < t.xoff[q + 1]; ++x, ++p) {
        const i64 c = t.xcol[x];
        if (c < 0 || c >= n) return pre + "external column out of range";
        if (block_of[c] == b) return pre + "external entry inside its own block";
        if (c != rows_col[p] || t.xsrc[x] != rows_src[p]) return pre + "external entry out of order";
        if (t.xu[x] >= nu || t.ucol[ub + t.xu[x]] != c)
          return pre + "external entry names the wrong distinct column";
        if (t.xsrc[x] < 0 || t.xsrc[x] >= nnz) return pre + "external source out of range";
        if (pos[block_of[c]] >= pos[b]) return pre + "processing order is not topological";
      }
      for (i64 x = t.ioff[q]; x < t.ioff[q + 1]; ++x, ++p) {
        const i64 kk = t.ik[x];
        if (kk < 0 || kk >= k) return pre + "internal entry not before its row";
        if (t.brow[s + kk] != rows_col[p] || t.isrc[x] != rows_src[p])
          return pre + "internal entry out of order";
        if (t.isrc[x] < 0 || t.isrc[x] >= nnz) return pre + "internal source out of range";
      }
    }
  }
  return std::string();
}

}  // namespace

static std::string check_ic_and_sell(const LowerPattern& lp, const DeviceLayouts& d);

// The checks of the IC(0) order, the sliced ELL and the two triangular-solve layouts are
// independent: the two solves' run on threads of their own, and the first failure in that
// order is reported.
std::string check_layouts(const LowerPattern& lp, const DeviceLayouts& d, const TriCaps& caps) {
  const i64 n = lp.n(), nnz = lp.nnz();
  if (d.n != n || d.nnz != nnz || (i64)d.off32.size() != n + 1 || (i64)d.col32.size() != nnz)
    return "CSR copy of the wrong size";
  const i64 nf = nnz - n;
  std::string err_fw, err_bw;
  ThreadGroup checks;
  checks.run([&] {
    std::vector<int32_t> foff(n + 1), fcol(nf), fsrc(nf);
    for (i64 i = 0; i <= n; ++i) foff[i] = (int32_t)(lp.offsets[i] - i);
    for (i64 i = 0; i < n; ++i)
      for (i64 p = lp.offsets[i], o = foff[i]; p < lp.offsets[i + 1] - 1; ++p, ++o) {
        fcol[o] = lp.cols[p];
        fsrc[o] = (int32_t)p;
      }
    err_fw = check_tri("forward", d.fw, n, nnz, caps, foff, fcol, fsrc);
    if (!err_fw.empty()) return;
    for (i64 q = 0; q < n; ++q)
      if (d.fw.dsrc[q] != lp.offsets[d.fw.brow[q] + 1] - 1) {
        err_fw = "forward layout: wrong diagonal";
        return;
      }
  });
  checks.run([&] {
    // backward rows, restated independently of device_layouts: entry (i, j) of L, i > j, goes
    // to logical row n-1-j; visiting i downwards appends each row's entries in ascending
    // logical column n-1-i
    std::vector<int32_t> boff(n + 1, 0), bcol(nf), bsrc(nf);
    for (i64 i = 0; i < n; ++i)
      for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p) boff[n - lp.cols[p]] += 1;
    for (i64 r = 0; r < n; ++r) boff[r + 1] += boff[r];
    {
      std::vector<int32_t> fill(boff.begin(), boff.end() - 1);
      for (i64 i = n - 1; i >= 0; --i)
        for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p) {
          const i64 q = fill[n - 1 - lp.cols[p]]++;
          bcol[q] = (int32_t)(n - 1 - i);
          bsrc[q] = (int32_t)p;
        }
    }
    err_bw = check_tri("backward", d.bw, n, nnz, caps, boff, bcol, bsrc);
    if (!err_bw.empty()) return;
    for (i64 q = 0; q < n; ++q)
      if (d.bw.dsrc[q] != lp.offsets[n - d.bw.brow[q]] - 1) {
        err_bw = "backward layout: wrong diagonal";
        return;
      }
  });
  std::string err = check_ic_and_sell(lp, d);
  checks.join();
  if (!err.empty()) return err;
  return err_fw.empty() ? err_bw : err_fw;
}

// The IC(0) order, the sliced ELL and its 16-bit and grouped column forms.
static std::string check_ic_and_sell(const LowerPattern& lp, const DeviceLayouts& d) {
  const i64 n = lp.n(), nnz = lp.nnz();
  // IC(0) order: a permutation in which every row comes after the rows it reads
  {
    if ((i64)d.ic_order.size() != n) return "IC(0) order of the wrong length";
    std::vector<int32_t> at(n, -1);
    for (i64 k = 0; k < n; ++k) {
      const i64 i = d.ic_order[k];
      if (i < 0 || i >= n || at[i] >= 0) return "IC(0) order is not a permutation";
      at[i] = (int32_t)k;
    }
    for (i64 i = 0; i < n; ++i)
      for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p)
        if (at[lp.cols[p]] >= at[i]) return "IC(0) order is not topological";
  }
  // sliced ELL
  if (d.slices != (n + 63) / 64 || (i64)d.soff.size() != d.slices + 1 || d.soff[0] != 0 ||
      (size_t)d.soff[d.slices] != d.scol.size() || d.scol.size() != d.ssrc.size() ||
      (i64)d.slen.size() != n)
    return "sliced-ELL arrays of the wrong size";
  for (i64 sl = 0; sl < d.slices; ++sl) {
    const i64 w = (i64)d.soff[sl + 1] - d.soff[sl];
    if (w < 64 || w % 64) return "sliced-ELL slice of a bad width";
  }
  for (i64 j = 0; j < n; ++j) {
    const i64 sl = j >> 6, width = ((i64)d.soff[sl + 1] - d.soff[sl]) / 64;
    if (d.slen[j] < 1 || d.slen[j] > width) return "sliced-ELL row longer than its slice";
  }
  for (size_t q = 0; q < d.scol.size(); ++q)
    if (d.scol[q] < 0 || d.scol[q] >= n || d.ssrc[q] < -1 || d.ssrc[q] >= nnz)
      return "sliced-ELL entry out of range";
  if (!d.sdel.empty()) {
    if (d.sdel.size() != d.scol.size()) return "sliced-ELL 16-bit offsets of the wrong size";
    for (i64 j = 0; j < n; ++j) {
      const i64 sl = j >> 6, base = d.soff[sl] + (j & 63);
      for (i64 k = 0; k < d.slen[j]; ++k)
        if (j + d.sdel[base + (k << 6)] != d.scol[base + (k << 6)])
          return "sliced-ELL 16-bit offset disagrees with its column";
    }
  }
  if (!d.ginfo.empty()) {
    // every row's grouped stream decodes to its sliced-ELL columns, in order
    if ((i64)d.goff.size() != d.slices + 1 || (i64)d.gnum.size() != d.slices ||
        (i64)d.ginfo.size() != n || d.goff[0] != 0 || (size_t)d.goff[d.slices] != d.gdel.size())
      return "grouped columns of the wrong size";
    for (i64 sl = 0; sl < d.slices; ++sl) {
      const i64 G = d.gnum[sl], span = (i64)d.goff[sl + 1] - d.goff[sl];
      if (G < 1 || G > 64 || span < 0 || span % G) return "grouped columns: bad slice";
    }
    for (i64 j = 0; j < n; ++j) {
      const uint32_t gi = d.ginfo[j];
      const i64 len = gi & 255, pos = (gi >> 8) & 255, g = (gi >> 16) & 63, sub = (gi >> 24) & 63;
      const i64 sl = j >> 6, G = d.gnum[sl], W = ((i64)d.goff[sl + 1] - d.goff[sl]) / G;
      if (len != d.slen[j] || pos >= len || g >= G || len > W || sub > (j & 63))
        return "grouped columns: bad row fields";
      const i64 base = d.soff[sl] + (j & 63);
      if (j - sub + d.gdel[d.goff[sl] + pos * G + g] != j)
        return "grouped columns: the row's own position does not hold its column";
      for (i64 k = 1; k < len; ++k) {
        const i64 q = k - 1 < pos ? k - 1 : k;
        if (j - sub + d.gdel[d.goff[sl] + q * G + g] != d.scol[base + (k << 6)])
          return "grouped columns disagree with the sliced ELL";
      }
    }
  }
  return std::string();
}

namespace {

// Sliced ELL of rows whose entries are listed by `row_entries(i, put)` (put(col, src) in the
// row's order).
template <class F>
std::string sweep_sell(i64 n, const std::vector<i64>& len, F row_entries, SweepLayout* out) {
  SweepLayout& s = *out;
  s.slices = (n + 63) / 64;
  s.soff.assign(s.slices + 1, 0);
  std::vector<i64> off(s.slices + 1, 0);
  for (i64 sl = 0; sl < s.slices; ++sl) {
    i64 w = 1;  // at least one column per slice, so that every slice has a valid base
    for (i64 j = sl * 64; j < std::min(n, (sl + 1) * 64); ++j) w = std::max(w, len[j]);
    off[sl + 1] = off[sl] + w * 64;
  }
  if (off[s.slices] >= (i64)INT32_MAX) return "strict triangle too large for 32-bit offsets";
  s.soff.assign(off.begin(), off.end());
  s.scol.assign(off[s.slices], 0);
  s.ssrc.assign(off[s.slices], -1);
  s.slen.resize(n);
  parallel_ranges(n, [&](i64 b, i64 e) {
    for (i64 i = b; i < e; ++i) {
      const i64 base = off[i >> 6] + (i & 63), w = (off[(i >> 6) + 1] - off[i >> 6]) / 64;
      i64 k = 0;
      row_entries(i, [&](i64 col, i64 src) {
        s.scol[base + (k << 6)] = (int32_t)col;
        s.ssrc[base + (k << 6)] = (int32_t)src;
        ++k;
      });
      s.slen[i] = (int32_t)k;
      for (; k < w; ++k) s.scol[base + (k << 6)] = (int32_t)i;
    }
  });
  return std::string();
}

}  // namespace

std::string sweep_layouts(const LowerPattern& lp, SweepLayout* lower, SweepLayout* upper) {
  const i64 n = lp.n();
  if (lp.nnz() >= (i64)INT32_MAX) return "factor too large for 32-bit sources";
  std::vector<i64> len(n);
  for (i64 i = 0; i < n; ++i) len[i] = lp.offsets[i + 1] - lp.offsets[i] - 1;
  std::string err = sweep_sell(n, len, [&](i64 i, auto put) {
    for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p) put(lp.cols[p], p);
  }, lower);
  if (!err.empty()) return err;
  TransposeLayout t = transpose_layout(lp);
  for (i64 j = 0; j < n; ++j) len[j] = t.offsets[j + 1] - t.offsets[j];
  err = sweep_sell(n, len, [&](i64 j, auto put) {
    for (i64 q = t.offsets[j]; q < t.offsets[j + 1]; ++q) put(t.rows[q], t.src[q]);
  }, upper);
  if (!err.empty()) return err;
  for (SweepLayout* s : {lower, upper}) {
    s->dsrc.resize(n);
    for (i64 i = 0; i < n; ++i) s->dsrc[i] = (int32_t)(lp.offsets[i + 1] - 1);
    // 16-bit distances from the row, when every one fits
    bool fits = true;
    for (i64 i = 0; i < n && fits; ++i) {
      const i64 base = s->soff[i >> 6] + (i & 63);
      for (i64 k = 0; k < s->slen[i]; ++k)
        if (std::abs((i64)s->scol[base + (k << 6)] - i) > 65535) { fits = false; break; }
    }
    s->sdelta.clear();
    if (fits) {
      s->sdelta.assign(s->scol.size(), 0);
      parallel_ranges(n, [&](i64 b, i64 e) {
        for (i64 i = b; i < e; ++i) {
          const i64 base = s->soff[i >> 6] + (i & 63);
          for (i64 k = 0; k < s->slen[i]; ++k)
            s->sdelta[base + (k << 6)] = (uint16_t)std::abs((i64)s->scol[base + (k << 6)] - i);
        }
      });
    }
  }
  return std::string();
}

std::string check_sweep_layouts(const LowerPattern& lp, const SweepLayout& lower,
                                const SweepLayout& upper) {
  const i64 n = lp.n(), nnz = lp.nnz();
  // upper rows restated: column j of L, ascending row
  std::vector<std::vector<std::pair<i64, i64>>> cols(n);
  for (i64 i = 0; i < n; ++i)
    for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p) cols[lp.cols[p]].push_back({i, p});
  for (int u = 0; u < 2; ++u) {
    const SweepLayout& s = u ? upper : lower;
    if (s.slices != (n + 63) / 64 || (i64)s.soff.size() != s.slices + 1 || s.soff[0] != 0 ||
        (size_t)s.soff[s.slices] != s.scol.size() || s.scol.size() != s.ssrc.size() ||
        (i64)s.slen.size() != n || (i64)s.dsrc.size() != n)
      return "sweep layout arrays of the wrong size";
    for (i64 sl = 0; sl < s.slices; ++sl) {
      const i64 w = (i64)s.soff[sl + 1] - s.soff[sl];
      if (w < 64 || w % 64) return "sweep layout slice of a bad width";
    }
    for (size_t q = 0; q < s.scol.size(); ++q)
      if (s.scol[q] < 0 || s.scol[q] >= n || s.ssrc[q] < -1 || s.ssrc[q] >= nnz)
        return "sweep layout entry out of range";
    for (i64 i = 0; i < n; ++i) {
      const i64 sl = i >> 6, base = s.soff[sl] + (i & 63), w = ((i64)s.soff[sl + 1] - s.soff[sl]) / 64;
      if (s.slen[i] < 0 || s.slen[i] > w) return "sweep layout row longer than its slice";
      if (s.dsrc[i] != lp.offsets[i + 1] - 1) return "sweep layout: wrong diagonal";
      const i64 want = u ? (i64)cols[i].size() : lp.offsets[i + 1] - lp.offsets[i] - 1;
      if (s.slen[i] != want) return "sweep layout row of the wrong length";
      for (i64 k = 0; k < s.slen[i]; ++k) {
        const i64 at = base + (k << 6);
        const i64 c = u ? cols[i][k].first : lp.cols[lp.offsets[i] + k];
        const i64 src = u ? cols[i][k].second : lp.offsets[i] + k;
        if (s.scol[at] != c || s.ssrc[at] != src) return "sweep layout entry out of order";
        if (!s.sdelta.empty() && (u ? i + s.sdelta[at] : i - s.sdelta[at]) != c)
          return "sweep layout: wrong 16-bit distance";
      }
    }
    if (!s.sdelta.empty() && s.sdelta.size() != s.scol.size())
      return "sweep layout: 16-bit distances of the wrong size";
  }
  return std::string();
}

namespace {

// Position of entry q of row i in a fused layout (sub-block i / 256, lane i % 256).
inline i64 fused_at(const FusedSweepLayout& f, i64 i, i64 q) {
  return f.eoff[i / FUSED_LANES] + (q >> 2) * 4 * FUSED_LANES + 4 * (i % FUSED_LANES) + (q & 3);
}

// Rows listed by `row_entries(i, put)` (put(col, src), ascending column) with lengths `len`,
// in blocks of R rows.
template <class F>
std::string fused_layout(i64 n, i64 R, const std::vector<i64>& len, F row_entries,
                         FusedSweepLayout* out) {
  const i64 L = FUSED_LANES;
  FusedSweepLayout& f = *out;
  f.rows = (int)R;
  f.blocks = (n + R - 1) / R;
  const i64 subs = (n + L - 1) / L;
  f.width = 0;
  for (i64 i = 0; i < n; ++i) f.width = std::max<int>(f.width, (int)len[i]);
  if (f.width > FUSED_WIDTH) return "row too long for the fused sweep kernel";
  f.elen.resize(n);
  for (i64 i = 0; i < n; ++i) f.elen[i] = (uint8_t)len[i];
  std::vector<i64> eoff(subs + 1, 0);
  for (i64 s = 0; s < subs; ++s) {
    i64 w = 0;
    for (i64 i = s * L; i < std::min(n, (s + 1) * L); ++i) w = std::max(w, len[i]);
    eoff[s + 1] = eoff[s] + ((w + 3) / 4) * 4 * L;
  }
  if (eoff[subs] >= (i64)INT32_MAX) return "strict triangle too large for 32-bit offsets";
  // each block's distinct columns outside it, ascending
  std::vector<std::vector<int32_t>> ext(f.blocks);
  parallel_ranges(f.blocks, [&](i64 b0, i64 b1) {
    for (i64 b = b0; b < b1; ++b) {
      std::vector<int32_t>& x = ext[b];
      for (i64 i = b * R; i < std::min(n, (b + 1) * R); ++i)
        row_entries(i, [&](i64 c, i64) {
          if (c < b * R || c >= (b + 1) * R) x.push_back((int32_t)c);
        });
      std::sort(x.begin(), x.end());
      x.erase(std::unique(x.begin(), x.end()), x.end());
    }
  });
  f.xoff.assign(f.blocks + 1, 0);
  f.max_ext = 0;
  for (i64 b = 0; b < f.blocks; ++b) {
    f.max_ext = std::max<int>(f.max_ext, (int)ext[b].size());
    f.xoff[b + 1] = f.xoff[b] + (int32_t)ext[b].size();
  }
  if (f.max_ext > FUSED_EXT) return "a block reads too many external columns for the fused sweep kernel";
  f.xcol.resize(f.xoff[f.blocks]);
  f.eoff.assign(eoff.begin(), eoff.end());
  f.esrc.assign(eoff[subs], -1);
  f.eidx.assign(eoff[subs], 0);
  parallel_ranges(f.blocks, [&](i64 b0, i64 b1) {
    for (i64 b = b0; b < b1; ++b) {
      const std::vector<int32_t>& x = ext[b];
      std::copy(x.begin(), x.end(), f.xcol.begin() + f.xoff[b]);
      for (i64 i = b * R; i < std::min(n, (b + 1) * R); ++i) {
        const i64 s = i / L, w = (eoff[s + 1] - eoff[s]) / L;
        i64 q = 0;
        row_entries(i, [&](i64 c, i64 src) {
          const i64 at = fused_at(f, i, q++);
          f.esrc[at] = (int32_t)src;
          if (c >= b * R && c < (b + 1) * R)
            f.eidx[at] = (uint16_t)(c - b * R);
          else
            f.eidx[at] = (uint16_t)(R + (std::lower_bound(x.begin(), x.end(), (int32_t)c) - x.begin()));
        });
        for (; q < w; ++q) f.eidx[fused_at(f, i, q)] = (uint16_t)(i - b * R);
      }
    }
  });
  return std::string();
}

}  // namespace

std::string fused_sweep_layouts(const LowerPattern& lp, int rows, FusedSweepLayout* lower,
                                FusedSweepLayout* upper) {
  const i64 n = lp.n();
  if (rows != FUSED_LANES && rows != 2 * FUSED_LANES) return "fused sweep blocks of 256 or 512 rows";
  if (lp.nnz() >= (i64)INT32_MAX) return "factor too large for 32-bit sources";
  std::vector<i64> len(n);
  for (i64 i = 0; i < n; ++i) len[i] = lp.offsets[i + 1] - lp.offsets[i] - 1;
  std::string err = fused_layout(n, rows, len, [&](i64 i, auto put) {
    for (i64 p = lp.offsets[i]; p < lp.offsets[i + 1] - 1; ++p) put(lp.cols[p], p);
  }, lower);
  if (!err.empty()) return err;
  TransposeLayout t = transpose_layout(lp);
  for (i64 j = 0; j < n; ++j) len[j] = t.offsets[j + 1] - t.offsets[j];
  return fused_layout(n, rows, len, [&](i64 j, auto put) {
    for (i64 q = t.offsets[j]; q < t.offsets[j + 1]; ++q) put(t.rows[q], t.src[q]);
  }, upper);
}

std::string check_fused_sweep_layouts(const LowerPattern& lp, const FusedSweepLayout& lower,
                                      const FusedSweepLayout& upper) {
  const i64 n = lp.n(), nnz = lp.nnz(), L = FUSED_LANES;
  const i64 subs = (n + L - 1) / L;
  TransposeLayout tr = transpose_layout(lp);
  for (int u = 0; u < 2; ++u) {
    const FusedSweepLayout& f = u ? upper : lower;
    const i64 R = f.rows;
    if ((R != L && R != 2 * L) || f.blocks != (n + R - 1) / R ||
        (i64)f.eoff.size() != subs + 1 || f.eoff[0] != 0 ||
        (size_t)f.eoff[subs] != f.esrc.size() || f.esrc.size() != f.eidx.size() ||
        (i64)f.elen.size() != n || (i64)f.xoff.size() != f.blocks + 1 || f.xoff[0] != 0 ||
        (size_t)f.xoff[f.blocks] != f.xcol.size())
      return "fused sweep layout arrays of the wrong size";
    for (i64 s = 0; s < subs; ++s) {
      const i64 span = (i64)f.eoff[s + 1] - f.eoff[s];
      if (span < 0 || span % (4 * L)) return "fused sweep sub-block of a bad width";
    }
    for (i64 b = 0; b < f.blocks; ++b) {
      const i64 x0 = f.xoff[b], nx = (i64)f.xoff[b + 1] - x0;
      if (nx < 0 || nx > FUSED_EXT) return "fused sweep block with too many external columns";
      for (i64 e = 0; e < nx; ++e) {
        const i64 c = f.xcol[x0 + e];
        if (c < 0 || c >= n || (e > 0 && c <= f.xcol[x0 + e - 1]))
          return "fused sweep external columns not ascending and distinct";
        // a lower block reads earlier blocks only, an upper one later blocks only: both are
        // processed before it
        if (u ? c < (b + 1) * R : c >= b * R) return "fused sweep external column in a block processed later";
      }
      for (i64 i = b * R; i < std::min(n, (b + 1) * R); ++i) {
        const i64 len = f.elen[i], span = (i64)f.eoff[i / L + 1] - f.eoff[i / L];
        const i64 want = u ? tr.offsets[i + 1] - tr.offsets[i] : lp.offsets[i + 1] - lp.offsets[i] - 1;
        if (len != want || 4 * L * ((len + 3) / 4) > span) return "fused sweep row of the wrong length";
        for (i64 q = 0; q < span / L; ++q) {
          const i64 at = fused_at(f, i, q);
          if (q >= len) {
            if (f.esrc[at] != -1 || f.eidx[at] != i - b * R) return "fused sweep padding not neutral";
            continue;
          }
          const i64 c = u ? tr.rows[tr.offsets[i] + q] : lp.cols[lp.offsets[i] + q];
          const i64 src = u ? tr.src[tr.offsets[i] + q] : lp.offsets[i] + q;
          if (f.esrc[at] != src || src < 0 || src >= nnz) return "fused sweep entry out of order";
          const i64 sl = f.eidx[at];
          const bool own = c >= b * R && c < (b + 1) * R;
          if (own ? sl != c - b * R : (sl < R || sl - R >= nx || f.xcol[x0 + sl - R] != c))
            return "fused sweep entry in the wrong LDS slot";
          if (u ? c <= i : c >= i) return "fused sweep entry outside its triangle";
        }
      }
    }
  }
  return std::string();
}

void fused_sweeps_host(const LowerPattern& lp, const FusedSweepLayout& lower,
                       const FusedSweepLayout& upper, const double* l, const double* r, int k,
                       double* z) {
  const i64 n = lp.n();
  std::vector<double> rd(n), y(n), prev(n), next(n);
  for (i64 i = 0; i < n; ++i) rd[i] = 1.0 / (double)(float)l[lp.offsets[i + 1] - 1];
  auto system = [&](const FusedSweepLayout& f, const double* rhs, double* out) {
    const i64 R = f.rows;
    for (i64 i = 0; i < n; ++i) prev[i] = rhs[i] * rd[i];
    std::vector<double> buf(R + FUSED_EXT);
    for (int j = 1; j <= k; ++j) {
      for (i64 b = 0; b < f.blocks; ++b) {
        // the kernel's LDS block: own rows, then the external columns, of the previous level
        for (i64 t = 0; t < R; ++t) buf[t] = b * R + t < n ? prev[b * R + t] : 0.0;
        for (i64 e = f.xoff[b]; e < f.xoff[b + 1]; ++e) buf[R + e - f.xoff[b]] = prev[f.xcol[e]];
        for (i64 i = b * R; i < std::min(n, (b + 1) * R); ++i) {
          double s = 0.0;
          for (i64 q = 0; q < f.elen[i]; ++q) {
            const i64 at = fused_at(f, i, q);
            s = s + (double)(float)l[f.esrc[at]] * buf[f.eidx[at]];
          }
          next[i] = (rhs[i] - s) * rd[i];
        }
      }
      // the iterates the kernel exchanges between sweeps: low 8 significand bits cleared
      if (j < k)
        for (i64 i = 0; i < n; ++i) {
          uint64_t u;
          std::memcpy(&u, &next[i], 8);
          u &= ~uint64_t(0xff);
          std::memcpy(&next[i], &u, 8);
        }
      prev.swap(next);
    }
    std::copy(prev.begin(), prev.end(), out);
  };
  system(lower, r, y.data());
  system(upper, y.data(), z);
}

void layout_precondition(const DeviceLayouts& d, const double* l, const double* r, double* z) {
  const i64 n = d.n;
  std::vector<double> y(n);
  // forward: y_i = (r_i - sum l_ij y_j) / l_ii, the sum in ascending j
  for (int32_t b : d.fw.border) {
    const i64 s = d.fw.bstart[b];
    for (i64 q = s; q < d.fw.bstart[b + 1]; ++q) {
      double acc = 0.0;
      for (i64 x = d.fw.xoff[q]; x < d.fw.xoff[q + 1]; ++x) acc = acc + l[d.fw.xsrc[x]] * y[d.fw.xcol[x]];
      for (i64 x = d.fw.ioff[q]; x < d.fw.ioff[q + 1]; ++x)
        acc = acc + l[d.fw.isrc[x]] * y[d.fw.brow[s + d.fw.ik[x]]];
      const i64 i = d.fw.brow[q];
      y[i] = (r[i] - acc) / l[d.fw.dsrc[q]];
    }
  }
  // backward in logical rows, vectors indexed physically (n-1-logical)
  for (int32_t b : d.bw.border) {
    const i64 s = d.bw.bstart[b];
    for (i64 q = s; q < d.bw.bstart[b + 1]; ++q) {
      const i64 lr = d.bw.brow[q];
      double acc = y[n - 1 - lr];
      for (i64 x = d.bw.xoff[q]; x < d.bw.xoff[q + 1]; ++x)
        acc = acc - l[d.bw.xsrc[x]] * z[n - 1 - d.bw.xcol[x]];
      for (i64 x = d.bw.ioff[q]; x < d.bw.ioff[q + 1]; ++x)
        acc = acc - l[d.bw.isrc[x]] * z[n - 1 - d.bw.brow[s + d.bw.ik[x]]];
      z[n - 1 - lr] = acc / l[d.bw.dsrc[q]];
    }
  }
}

}  // namespace ml
