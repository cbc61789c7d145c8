// Gmsh 4.1 ASCII reader: the parse and the set rules of maille-io/src/msh/reader.rs:23-471 and
// the numbering of numbering.rs:11-99, as a single-pass native reader.
//
// Differences of route, not of result:
//  - sections are located by one structure pass and then parsed straight from the text through
//    cursors over their byte ranges (no per-line allocation);
//  - node tags go to a dense index when they are small (Gmsh writes 1..N), else to a hash map;
//  - boundary faces and edges are found by looking up only the keys the physical groups ask
//    for during one scan of the cells in (element, face) order, which gives the first match,
//    the element the reference's full face map keeps (reader.rs:363-385);
//  - whitespace is ASCII whitespace (Rust's trim and split_whitespace also take the Unicode
//    spaces, which Gmsh never writes).
#include "msh.hpp"

#include <algorithm>
#include <clocale>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <locale.h>
#include <unordered_map>

namespace ml {
namespace msh {
namespace {

inline bool is_ws(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
}

struct Line {
  int64_t no;     // 1-based line number
  const char* r;  // raw start of the line
  const char* b;  // trimmed content [b, e)
  const char* e;
};

// The non-empty trimmed lines of [p, end); `no` is the number of the line starting at p.
struct Lines {
  const char* p;
  const char* end;
  int64_t no;
  bool next(Line& out) {
    while (p < end) {
      const char* nl = (const char*)memchr(p, '\n', (size_t)(end - p));
      const char* le = nl ? nl : end;
      const char* b = p;
      const char* e = le;
      while (b < e && is_ws(*b)) ++b;
      while (e > b && is_ws(e[-1])) --e;
      Line l{no, p, b, e};
      p = nl ? nl + 1 : end;
      ++no;
      if (b < e) {
        out = l;
        return true;
      }
    }
    return false;
  }
};

struct Section {
  const char* b = nullptr;  // body [b, e)
  const char* e = nullptr;
  int64_t first_no = 0;  // number of the line at b
  int64_t last_no = 0;   // last non-empty body line, 0 for an empty body
};

Error syntax(int64_t line, const std::string& m) { return Error(E_SYNTAX, line, m); }

// reader.rs:46-80
std::map<std::string, Section> split_sections(const char* text, size_t len) {
  std::map<std::string, Section> sections;
  Lines it{text, text + len, 1};
  Line l;
  bool open = false;
  std::string name;
  Section cur;
  int64_t last = 0;
  while (it.next(l)) {
    last = l.no;
    bool dollar = *l.b == '$';
    if (!open) {
      if (dollar && !(l.e - l.b >= 4 && memcmp(l.b + 1, "End", 3) == 0)) {
        open = true;
        name.assign(l.b + 1, l.e);
        cur = Section();
        cur.b = it.p;
        cur.first_no = it.no;
      }
      continue;
    }
    if (dollar) {
      std::string other(l.b + 1, l.e);
      if (other.size() == name.size() + 3 && other.compare(0, 3, "End") == 0 &&
          other.compare(3, std::string::npos, name) == 0) {
        cur.e = l.r;
        sections[name] = cur;
        open = false;
      } else {
        throw syntax(l.no, "'$" + other + "' inside section $" + name + ": $End" + name +
                               " is missing");
      }
    } else {
      cur.last_no = l.no;
    }
  }
  if (open)
    throw syntax(last, "section $" + name + " is not closed: $End" + name + " is missing");
  return sections;
}

struct Tok {
  const char* b;
  const char* e;
};

void split(const Line& l, std::vector<Tok>& out) {
  out.clear();
  const char* p = l.b;
  while (p < l.e) {
    while (p < l.e && is_ws(*p)) ++p;
    if (p >= l.e) break;
    const char* s = p;
    while (p < l.e && !is_ws(*p)) ++p;
    out.push_back({s, p});
  }
}

// reader.rs:83-114
struct Cursor {
  std::string name;
  Lines it;
  int64_t last_no;
  std::vector<Tok> toks;
  Cursor(const char* n, const Section& s) : name(n), it{s.b, s.e, s.first_no}, last_no(s.last_no) {}
  Line next_line() {
    Line l;
    if (!it.next(l)) throw syntax(last_no, "section $" + name + " ends too early");
    return l;
  }
  Line next() {
    Line l = next_line();
    split(l, toks);
    return l;
  }
};

const Tok& tok_at(int64_t line, const std::vector<Tok>& t, size_t i, const char* what) {
  if (i >= t.size()) throw syntax(line, std::string(what) + " is missing");
  return t[i];
}

[[noreturn]] void bad_number(int64_t line, const char* what, const Tok& t) {
  throw syntax(line, std::string(what) + ": '" + std::string(t.b, t.e) + "' is not a valid number");
}

// Rust's FromStr for i64 / usize / u32: optional sign ('+' only when unsigned), decimal digits,
// no overflow (reader.rs:116-126).
uint64_t parse_digits(int64_t line, const char* what, const Tok& t, const char* p, uint64_t lim) {
  if (p == t.e) bad_number(line, what, t);
  uint64_t v = 0;
  for (; p < t.e; ++p) {
    unsigned d = (unsigned)(*p - '0');
    if (d > 9 || v > (lim - d) / 10) bad_number(line, what, t);
    v = v * 10 + d;
  }
  return v;
}

int64_t parse_i64(int64_t line, const std::vector<Tok>& toks, size_t i, const char* what) {
  const Tok& t = tok_at(line, toks, i, what);
  const char* p = t.b;
  bool neg = false;
  if (*p == '+' || *p == '-') neg = *p++ == '-';
  uint64_t v = parse_digits(line, what, t, p, neg ? (uint64_t)INT64_MAX + 1 : (uint64_t)INT64_MAX);
  return neg ? (int64_t)(0 - v) : (int64_t)v;
}

uint64_t parse_unsigned(int64_t line, const std::vector<Tok>& toks, size_t i, const char* what,
                        uint64_t lim = UINT64_MAX) {
  const Tok& t = tok_at(line, toks, i, what);
  const char* p = t.b;
  if (*p == '+') ++p;
  return parse_digits(line, what, t, p, lim);
}

bool word_ci(const char* p, const char* e, const char* w) {
  size_t n = strlen(w);
  if ((size_t)(e - p) != n) return false;
  for (size_t k = 0; k < n; ++k)
    if ((p[k] | 0x20) != w[k]) return false;
  return true;
}

locale_t c_locale() {
  static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
  return loc;
}

// Rust's FromStr for f64: [sign] (inf | infinity | nan, any case | digits [. digits] [exp] |
// . digits [exp]), correctly rounded, overflow to infinity.
double parse_f64(int64_t line, const std::vector<Tok>& toks, size_t i, const char* what) {
  const Tok& t = tok_at(line, toks, i, what);
  const char* p = t.b;
  if (*p == '+' || *p == '-') ++p;
  bool ok;
  if (word_ci(p, t.e, "inf") || word_ci(p, t.e, "infinity") || word_ci(p, t.e, "nan")) {
    ok = true;
  } else {
    const char* q = p;
    size_t d1 = 0, d2 = 0;
    while (q < t.e && (unsigned)(*q - '0') <= 9) ++q, ++d1;
    if (q < t.e && *q == '.') {
      ++q;
      while (q < t.e && (unsigned)(*q - '0') <= 9) ++q, ++d2;
    }
    ok = d1 + d2 > 0;
    if (ok && q < t.e && (*q == 'e' || *q == 'E')) {
      ++q;
      if (q < t.e && (*q == '+' || *q == '-')) ++q;
      size_t d3 = 0;
      while (q < t.e && (unsigned)(*q - '0') <= 9) ++q, ++d3;
      ok = d3 > 0;
    }
    ok = ok && q == t.e;
  }
  if (!ok) bad_number(line, what, t);
  char small[64];
  std::string big;
  const size_t n = (size_t)(t.e - t.b);
  const char* s;
  if (n < sizeof(small)) {
    memcpy(small, t.b, n);
    small[n] = 0;
    s = small;
  } else {
    big.assign(t.b, t.e);
    s = big.c_str();
  }
  const char* src = (*s == '+') ? s + 1 : s;
  char* end = nullptr;
  double v = strtod_l(src, &end, c_locale());
  if (end != s + n) bad_number(line, what, t);
  return v;
}

// reader.rs:128-142
void check_format(const Section& s) {
  Cursor c("MeshFormat", s);
  Line l = c.next();
  std::string version = c.toks.empty() ? "" : std::string(c.toks[0].b, c.toks[0].e);
  if (version != "4.1")
    throw Error(E_UNSUPPORTED_VERSION, l.no,
                "format " + version + " is not supported; export with Mesh.MshFileVersion = 4.1");
  if (c.toks.size() < 2 || std::string(c.toks[1].b, c.toks[1].e) != "0")
    throw Error(E_BINARY, l.no, "binary file; export with Mesh.Binary = 0");
}

typedef std::pair<int64_t, int64_t> Entity;  // (dimension, tag)

// reader.rs:144-169
std::map<Entity, std::string> parse_names(const Section& s) {
  Cursor c("PhysicalNames", s);
  Line l = c.next();
  uint64_t count = parse_unsigned(l.no, c.toks, 0, "the number of physical groups");
  std::map<Entity, std::string> names;
  for (uint64_t k = 0; k < count; ++k) {
    Line t = c.next_line();
    const char* op = (const char*)memchr(t.b, '"', (size_t)(t.e - t.b));
    const char* cl = nullptr;
    for (const char* p = t.e; p > t.b; --p)
      if (p[-1] == '"') {
        cl = p - 1;
        break;
      }
    if (!op || !cl || cl <= op) throw syntax(t.no, "the group name must be in double quotes");
    Line head{t.no, t.r, t.b, op};
    std::vector<Tok> h;
    split(head, h);
    int64_t dim = (int64_t)parse_unsigned(t.no, h, 0, "the group dimension");
    int64_t tag = parse_i64(t.no, h, 1, "the group number");
    names[{dim, tag}] = std::string(op + 1, cl);
  }
  return names;
}

// reader.rs:172-199
std::map<Entity, std::vector<int64_t>> parse_entities(const Section& s) {
  Cursor c("Entities", s);
  Line l = c.next();
  uint64_t counts[4];
  for (int d = 0; d < 4; ++d) counts[d] = parse_unsigned(l.no, c.toks, d, "the number of entities");
  std::map<Entity, std::vector<int64_t>> entities;
  for (int d = 0; d < 4; ++d)
    for (uint64_t k = 0; k < counts[d]; ++k) {
      Line t = c.next();
      int64_t tag = parse_i64(t.no, c.toks, 0, "the entity number");
      size_t at = d == 0 ? 4 : 7;
      uint64_t n = parse_unsigned(t.no, c.toks, at, "the entity's number of physical groups");
      std::vector<int64_t> phys;
      for (uint64_t j = 0; j < n; ++j)
        phys.push_back(parse_i64(t.no, c.toks, at + 1 + j, "a physical group number"));
      entities[{d, tag}] = std::move(phys);
    }
  return entities;
}

// Gmsh tag -> node index: dense below a bound proportional to the announced count, else hashed.
struct NodeIndex {
  std::vector<int64_t> dense;
  std::unordered_map<int64_t, int64_t> sparse;
  int64_t dense_limit = 0;
  int64_t find(int64_t tag) const {
    if (tag >= 0 && tag < (int64_t)dense.size()) return dense[tag];
    if (tag >= 0 && tag < dense_limit) return -1;
    auto it = sparse.find(tag);
    return it == sparse.end() ? -1 : it->second;
  }
  void insert(int64_t tag, int64_t i) {
    if (tag >= 0 && tag < dense_limit) {
      if (tag >= (int64_t)dense.size()) dense.resize(std::max<size_t>(tag + 1, dense.size() * 2), -1);
      dense[tag] = i;
    } else {
      sparse[tag] = i;
    }
  }
};

// reader.rs:202-247
void parse_nodes(const Section& s, std::vector<double>& xyz, NodeIndex& index) {
  Cursor c("Nodes", s);
  Line h = c.next();
  uint64_t blocks = parse_unsigned(h.no, c.toks, 0, "the number of node blocks");
  uint64_t total = parse_unsigned(h.no, c.toks, 1, "the number of nodes");
  // a node takes at least 8 bytes of the section (tag line and coordinate line)
  const uint64_t most = (uint64_t)(s.e - s.b) / 8;
  index.dense_limit = (int64_t)std::min<uint64_t>(total, most) * 2 + 1024;
  std::vector<std::pair<int64_t, int64_t>> tags;  // (line, tag)
  for (uint64_t b = 0; b < blocks; ++b) {
    Line bl = c.next();
    uint64_t count = parse_unsigned(bl.no, c.toks, 3, "the number of nodes of the block");
    tags.clear();
    for (uint64_t k = 0; k < count; ++k) {
      Line t = c.next();
      tags.push_back({t.no, parse_i64(t.no, c.toks, 0, "the node number")});
    }
    for (const auto& tg : tags) {
      Line t = c.next();
      double x = parse_f64(t.no, c.toks, 0, "x");
      double y = parse_f64(t.no, c.toks, 1, "y");
      double z = parse_f64(t.no, c.toks, 2, "z");
      if (!(std::isfinite(x) && std::isfinite(y) && std::isfinite(z)))
        throw syntax(t.no, "a coordinate of node " + std::to_string(tg.second) +
                               " is not finite (NaN or infinite)");
      if (index.find(tg.second) >= 0)
        throw Error(E_DUPLICATE_NODE_TAG, tg.first,
                    "node " + std::to_string(tg.second) + " is defined twice");
      index.insert(tg.second, (int64_t)xyz.size() / 3);
      xyz.push_back(x);
      xyz.push_back(y);
      xyz.push_back(z);
    }
  }
  if (xyz.size() / 3 != total)
    throw syntax(h.no, std::to_string(total) + " nodes announced, " +
                           std::to_string(xyz.size() / 3) + " found");
}

struct Block {
  Entity entity;
  uint32_t type;
  int dim;
  int width;
  int64_t count;
  size_t node_off;   // into Elements::nodes
  size_t line_off;   // into Elements::lines (blocks of dimension <= 2)
  int64_t cell0 = -1;  // first cell index (blocks of the mesh dimension)
};

struct Elements {
  std::vector<Block> blocks;
  std::vector<int64_t> nodes;
  std::vector<int64_t> lines;
};

// reader.rs:250-261
bool element_info(uint32_t type, int& dim, int& width) {
  switch (type) {
    case 15: dim = 0; width = 1; return true;
    case 1: dim = 1; width = 2; return true;
    case 8: dim = 1; width = 3; return true;
    case 2: dim = 2; width = 3; return true;
    case 9: dim = 2; width = 6; return true;
    case 4: dim = 3; width = 4; return true;
    case 11: dim = 3; width = 10; return true;
    default: return false;
  }
}

const char* gmsh_name(uint32_t t) {
  switch (t) {
    case 3: return "quadrangle";
    case 5: return "hexahedron";
    case 6: return "prism";
    case 7: return "pyramid";
    case 10: return "quadrangle-9";
    case 12: return "hexahedron-27";
    case 16: return "quadrangle-8";
    case 17: return "hexahedron-20";
    default: return "other";
  }
}

// reader.rs:273-317
Elements parse_elements(const Section& s, const NodeIndex& index) {
  Cursor c("Elements", s);
  Line h = c.next();
  uint64_t nblocks = parse_unsigned(h.no, c.toks, 0, "the number of element blocks");
  Elements out;
  for (uint64_t k = 0; k < nblocks; ++k) {
    Line bl = c.next();
    Block b;
    b.entity.first = (int64_t)parse_unsigned(bl.no, c.toks, 0, "the entity dimension");
    b.entity.second = parse_i64(bl.no, c.toks, 1, "the entity number");
    b.type = (uint32_t)parse_unsigned(bl.no, c.toks, 2, "the element type", UINT32_MAX);
    uint64_t count = parse_unsigned(bl.no, c.toks, 3, "the number of elements of the block");
    if (!element_info(b.type, b.dim, b.width))
      throw Error(E_UNSUPPORTED_ELEMENT, bl.no,
                  "element type " + std::to_string(b.type) + " (" + gmsh_name(b.type) +
                      ") is not supported: mesh in triangles (Tri3, Tri6) or tetrahedra "
                      "(Tet4, Tet10)");
    b.node_off = out.nodes.size();
    b.line_off = out.lines.size();
    b.count = 0;
    for (uint64_t j = 0; j < count; ++j) {
      Line el = c.next();
      if (c.toks.size() != (size_t)b.width + 1)
        throw syntax(el.no, "an element of type " + std::to_string(b.type) + " must have " +
                                std::to_string(b.width) + " nodes, it has " +
                                std::to_string(c.toks.empty() ? 0 : c.toks.size() - 1));
      for (int q = 1; q <= b.width; ++q) {
        int64_t tag = parse_i64(el.no, c.toks, q, "a node number");
        int64_t i = index.find(tag);
        if (i < 0)
          throw Error(E_UNKNOWN_NODE, el.no,
                      "node " + std::to_string(tag) + " is not defined in $Nodes");
        out.nodes.push_back(i);
      }
      if (b.dim <= 2) out.lines.push_back(el.no);
      ++b.count;
    }
    out.blocks.push_back(b);
  }
  return out;
}

// shape.rs:129-135
const int TET_FACES[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {2, 3, 0}};
const int TET10_EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
const int TRI6_EDGES[3][2] = {{0, 1}, {1, 2}, {2, 0}};

int tet_mid(int a, int b) {
  for (int k = 0; k < 6; ++k)
    if ((TET10_EDGES[k][0] == a && TET10_EDGES[k][1] == b) ||
        (TET10_EDGES[k][0] == b && TET10_EDGES[k][1] == a))
      return 4 + k;
  return 4;
}

// numbering.rs:59-72: vertices first, then mid-nodes; returns the count
int face_positions(int nen, int f, int* pos) {
  switch (nen) {
    case 3: pos[0] = TRI6_EDGES[f][0]; pos[1] = TRI6_EDGES[f][1]; return 2;
    case 6: pos[0] = TRI6_EDGES[f][0]; pos[1] = TRI6_EDGES[f][1]; pos[2] = 3 + f; return 3;
    case 4: for (int k = 0; k < 3; ++k) pos[k] = TET_FACES[f][k]; return 3;
    default: {
      const int* v = TET_FACES[f];
      pos[0] = v[0]; pos[1] = v[1]; pos[2] = v[2];
      pos[3] = tet_mid(v[0], v[1]); pos[4] = tet_mid(v[1], v[2]); pos[5] = tet_mid(v[2], v[0]);
      return 6;
    }
  }
}

// numbering.rs:85-91 (tetrahedra; a triangle's edges are its faces)
int edge_positions(int nen, int k, int* pos) {
  pos[0] = TET10_EDGES[k][0];
  pos[1] = TET10_EDGES[k][1];
  if (nen == 10) {
    pos[2] = 4 + k;
    return 3;
  }
  return 2;
}

// numbering.rs:11-15
void swap_tet10_mids(int64_t* n) { std::swap(n[8], n[9]); }

// numbering.rs:20-46
void orient(int nen, int64_t* n, const double* xyz) {
  auto P = [&](int i, int c) { return xyz[3 * n[i] + c]; };
  bool reversed;
  if (nen == 3 || nen == 6) {
    reversed = (P(1, 0) - P(0, 0)) * (P(2, 1) - P(0, 1)) - (P(2, 0) - P(0, 0)) * (P(1, 1) - P(0, 1)) < 0.0;
  } else {
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = P(1, k) - P(0, k);
      b[k] = P(2, k) - P(0, k);
      c[k] = P(3, k) - P(0, k);
    }
    reversed = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                   a[2] * (b[0] * c[1] - b[1] * c[0]) <
               0.0;
  }
  if (!reversed) return;
  std::swap(n[0], n[1]);
  if (nen == 6) std::swap(n[4], n[5]);
  if (nen == 10) {
    std::swap(n[5], n[6]);
    std::swap(n[7], n[8]);
  }
}

// numbering.rs:94-99: sorted vertices, padded with -1 (a 2-vertex key never meets a 3-vertex one)
struct Key {
  int64_t v[3];
  bool operator==(const Key& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    uint64_t h = (uint64_t)k.v[0] * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)k.v[1] + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.v[2] + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};
Key make_key(const int64_t* v, int n) {
  Key k{{v[0], v[1], n > 2 ? v[2] : -1}};
  if (k.v[0] > k.v[1]) std::swap(k.v[0], k.v[1]);
  if (n > 2) {
    if (k.v[1] > k.v[2]) std::swap(k.v[1], k.v[2]);
    if (k.v[0] > k.v[1]) std::swap(k.v[0], k.v[1]);
  }
  return k;
}
int64_t key_min(const Key& k) { return k.v[0]; }

const char* kind_name(int nen) {
  switch (nen) {
    case 3: return "Tri3";
    case 6: return "Tri6";
    case 4: return "Tet4";
    default: return "Tet10";
  }
}

int cell_nen(uint32_t type) { return type == 2 ? 3 : type == 9 ? 6 : type == 4 ? 4 : 10; }

// reader.rs:461-471
void check_planar(const std::vector<double>& xyz) {
  double span = 0.0;
  for (size_t i = 0; i < xyz.size(); i += 3) span = std::fmax(std::fmax(span, std::fabs(xyz[i])), std::fabs(xyz[i + 1]));
  double tol = 1e-9 * std::fmax(span, 2.2250738585072014e-308);
  for (size_t i = 0; i < xyz.size(); i += 3)
    if (std::fabs(xyz[i + 2]) > tol)
      throw Error(E_NON_PLANAR, 0, "triangle mesh outside the plane z = 0 (z = " +
                                       std::to_string(xyz[i + 2]) + ")");
}

template <typename T>
void sort_dedup(std::vector<T>& v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// reader.rs:320-458
void assemble(File& f, Elements& el, const std::map<Entity, std::string>& names,
              const std::map<Entity, std::vector<int64_t>>& entities) {
  bool tets = false, tris = false;
  for (const Block& b : el.blocks) {
    tets |= b.type == 4 || b.type == 11;
    tris |= b.type == 2 || b.type == 9;
  }
  if (!tets && !tris)
    throw Error(E_NO_CELLS, 0, "the file holds no triangle and no tetrahedron");
  const int dim = tets ? 3 : 2;
  f.dim = dim;
  if (dim == 2) check_planar(f.xyz);

  int nen = 0;
  int64_t ncells = 0;
  for (Block& b : el.blocks) {
    if (b.dim != dim) continue;
    int k = cell_nen(b.type);
    if (nen && nen != k)
      throw Error(E_MIXED_CELL_KINDS, 0, std::string("the file mixes ") + kind_name(nen) +
                                             " and " + kind_name(k) + " elements: use one order");
    nen = k;
    b.cell0 = ncells;
    ncells += b.count;
  }
  f.nen = nen;
  f.cells.resize((size_t)ncells * nen);
  for (const Block& b : el.blocks) {
    if (b.dim != dim) continue;
    for (int64_t j = 0; j < b.count; ++j) {
      int64_t* cell = &f.cells[(size_t)(b.cell0 + j) * nen];
      std::memcpy(cell, &el.nodes[b.node_off + (size_t)j * nen], sizeof(int64_t) * nen);
      if (nen == 10) swap_tet10_mids(cell);
      orient(nen, cell, f.xyz.data());
    }
  }

  auto physical = [&](const Entity& e) -> const std::vector<int64_t>* {
    auto it = entities.find(e);
    return (it == entities.end() || it->second.empty()) ? nullptr : &it->second;
  };

  // Keys asked for by the boundary groups, then one scan of the cells in (element, face) order.
  const int64_t nn = f.n_nodes();
  std::unordered_map<Key, std::pair<int64_t, int64_t>, KeyHash> fq, eq;
  std::vector<uint8_t> mark;
  for (const Block& b : el.blocks) {
    if (!physical(b.entity)) continue;
    bool face = b.dim + 1 == dim, edge = !face && b.dim == 1;
    if (!face && !edge) continue;
    if (mark.empty()) mark.assign((size_t)nn, 0);
    for (int64_t j = 0; j < b.count; ++j) {
      Key k = make_key(&el.nodes[b.node_off + (size_t)j * b.width], face ? dim : 2);
      (face ? fq : eq).emplace(k, std::make_pair((int64_t)-1, (int64_t)-1));
      mark[key_min(k)] |= face ? 1 : 2;
    }
  }
  if (!fq.empty() || !eq.empty()) {
    const int nf = dim == 2 ? 3 : 4;
    int pos[6];
    int64_t v[3];
    for (int64_t e = 0; e < ncells; ++e) {
      const int64_t* cell = &f.cells[(size_t)e * nen];
      if (!fq.empty())
        for (int q = 0; q < nf; ++q) {
          face_positions(nen, q, pos);
          for (int k = 0; k < dim; ++k) v[k] = cell[pos[k]];
          Key key = make_key(v, dim);
          if (!(mark[key_min(key)] & 1)) continue;
          auto it = fq.find(key);
          if (it != fq.end() && it->second.first < 0) it->second = {e, q};
        }
      if (!eq.empty())
        for (int q = 0; q < 6; ++q) {
          v[0] = cell[TET10_EDGES[q][0]];
          v[1] = cell[TET10_EDGES[q][1]];
          Key key = make_key(v, 2);
          if (!(mark[key_min(key)] & 2)) continue;
          auto it = eq.find(key);
          if (it != eq.end() && it->second.first < 0) it->second = {e, q};
        }
    }
  }

  // cells of each entity, in block order
  std::map<Entity, std::vector<std::pair<int64_t, int64_t>>> entity_cells;
  for (const Block& b : el.blocks)
    if (b.dim == dim) entity_cells[b.entity].push_back({b.cell0, b.count});

  std::map<std::string, Entity> owner;
  for (const Block& b : el.blocks) {
    const std::vector<int64_t>* phys = physical(b.entity);
    if (!phys) continue;
    for (int64_t tag : *phys) {
      Entity group{b.entity.first, tag};
      auto nm = names.find(group);
      std::string name = nm != names.end() ? nm->second : std::to_string(tag);
      auto ow = owner.emplace(name, group);
      if (!ow.second && ow.first->second != group)
        throw Error(E_DUPLICATE_SET_NAME, 0,
                    "two physical groups are named '" + name + "': give them different names");
      Set& s = f.sets[name];
      int pos[6];
      if (b.dim == dim) {
        for (const auto& r : entity_cells[b.entity])
          for (int64_t c = r.first; c < r.first + r.second; ++c) s.elements.push_back(c);
      } else if (b.dim + 1 == dim || b.dim == 1) {
        const bool face = b.dim + 1 == dim;
        auto& q = face ? fq : eq;
        for (int64_t j = 0; j < b.count; ++j) {
          Key k = make_key(&el.nodes[b.node_off + (size_t)j * b.width], face ? dim : 2);
          auto it = q.find(k);
          if (it == q.end() || it->second.first < 0) {
            int64_t line = el.lines[b.line_off + (size_t)j];
            if (face)
              throw Error(E_BOUNDARY_FACE_NOT_FOUND, line,
                          "an element of group '" + name + "' is no element face: the boundary "
                          "must be meshed with the domain");
            throw Error(E_BOUNDARY_EDGE_NOT_FOUND, line,
                        "a line of group '" + name + "' is no element edge: the line must "
                        "follow mesh edges");
          }
          const int64_t e = it->second.first, loc = it->second.second;
          int np = face ? face_positions(nen, (int)loc, pos) : edge_positions(nen, (int)loc, pos);
          (face ? s.faces : s.edges).push_back({e, loc});
          for (int p = 0; p < np; ++p) s.nodes.push_back(f.cells[(size_t)e * nen + pos[p]]);
        }
      } else {
        const size_t n = (size_t)b.count * b.width;
        s.nodes.insert(s.nodes.end(), el.nodes.begin() + b.node_off,
                       el.nodes.begin() + b.node_off + n);
      }
    }
  }
  for (auto& kv : f.sets) {
    sort_dedup(kv.second.nodes);
    sort_dedup(kv.second.faces);
    sort_dedup(kv.second.edges);
    sort_dedup(kv.second.elements);
  }
}

}  // namespace

// reader.rs:23-39
File read(const char* text, size_t length) {
  auto sections = split_sections(text, length);
  auto section = [&](const char* name) -> const Section& {
    auto it = sections.find(name);
    if (it == sections.end())
      throw Error(E_MISSING_SECTION, 0, std::string("section $") + name + " is missing");
    return it->second;
  };
  check_format(section("MeshFormat"));
  std::map<Entity, std::string> names;
  std::map<Entity, std::vector<int64_t>> entities;
  if (sections.count("PhysicalNames")) names = parse_names(sections["PhysicalNames"]);
  if (sections.count("Entities")) entities = parse_entities(sections["Entities"]);
  File f;
  NodeIndex index;
  parse_nodes(section("Nodes"), f.xyz, index);
  Elements el = parse_elements(section("Elements"), index);
  assemble(f, el, names, entities);
  return f;
}

}  // namespace msh
}  // namespace ml
