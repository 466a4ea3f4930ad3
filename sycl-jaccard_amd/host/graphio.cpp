// Graph file formats and host-side graph preparation (see graphio.hpp).
#include "graphio.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <tuple>
#include <unordered_set>

namespace jcio {

namespace {

bool ends_with(const std::string& s, const char* suf) {
  const size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

void set_err(std::string* err, const std::string& msg) {
  if (err) *err = msg;
}

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

}  // namespace

int file_kind(const std::string& path) {
  if (ends_with(path, ".mtx")) return 0;
  if (ends_with(path, ".csr")) return 1;
  return -1;
}

int read_mtx(const std::string& path, Coo& out, bool drop_weights, std::string* err) {
  std::ifstream in(path);
  if (!in) {
    set_err(err, "cannot open " + path);
    return kErrOpen;
  }
  out = Coo{};
  std::string line;
  bool header_seen = false;
  // comment / banner lines
  while (in.peek() == '%') {
    std::getline(in, line);
    if (!header_seen && line.rfind("%%MatrixMarket", 0) == 0) {
      header_seen = true;
      std::istringstream hs(line.substr(14));
      std::string object, format, field, symmetry;
      hs >> object >> format >> field >> symmetry;
      object = lower(object), format = lower(format), field = lower(field), symmetry = lower(symmetry);
      if (object != "matrix") {
        set_err(err, "can only process matrix MTX files");
        return kErrNotMatrix;
      }
      if (format != "coordinate") {
        set_err(err, "can only process coordinate MTX files");
        return kErrNotCoordinate;
      }
      if (field == "complex") {
        set_err(err, "cannot process complex weights");
        return kErrComplex;
      }
      out.weighted = field == "real" || field == "double" || field == "integer";
      out.directed = symmetry == "general";
    }
  }
  int64_t rows = 0, cols = 0, nnz = 0;
  if (!(in >> rows >> cols >> nnz) || rows < 0 || cols < 0 || nnz < 0) {
    set_err(err, "bad MTX size line in " + path);
    return kErrFormat;
  }
  out.src.reserve((size_t)nnz);
  out.dst.reserve((size_t)nnz);
  out.w.reserve((size_t)nnz);
  int64_t maxid = 0;
  for (int64_t k = 0; k < nnz; ++k) {
    int64_t a, b;
    double v = 1.0;
    if (!(in >> a >> b)) break;
    if (out.weighted && !(in >> v)) break;
    if (a < 1 || b < 1 || a > INT32_MAX || b > INT32_MAX) {
      set_err(err, "MTX entry outside 1..2^31-1 in " + path);
      return kErrFormat;
    }
    out.src.push_back((int32_t)(a - 1));
    out.dst.push_back((int32_t)(b - 1));
    out.w.push_back(drop_weights ? 1.0f : (float)v);
    maxid = std::max(maxid, std::max(a, b));
  }
  out.n = std::max<int64_t>(std::max(rows, cols), maxid);
  return kOk;
}

int write_mtx(const std::string& path, const Coo& coo, bool weighted, bool directed, std::string* err) {
  std::ofstream os(path, std::ios::out | std::ios::trunc);
  if (!os) {
    set_err(err, "cannot open " + path);
    return kErrOpen;
  }
  os << "%%MatrixMarket matrix coordinate " << (weighted ? "real " : "pattern ")
     << (directed ? "general" : "symmetric") << "\n";
  os << coo.n << " " << coo.n << " " << coo.src.size() << "\n";
  // column-major order, as MatrixMarket files are usually written
  std::vector<size_t> ord(coo.src.size());
  std::iota(ord.begin(), ord.end(), (size_t)0);
  std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
    return std::tie(coo.dst[x], coo.src[x], coo.w[x]) < std::tie(coo.dst[y], coo.src[y], coo.w[y]);
  });
  for (size_t k : ord) {
    os << (int64_t)coo.src[k] + 1 << " " << (int64_t)coo.dst[k] + 1;
    if (weighted) os << " " << coo.w[k];  // default stream precision (6 significant digits)
    os << "\n";
  }
  return os ? kOk : kErrFormat;
}

int read_csr(const std::string& path, Csr& out, std::string* err) {
  std::ifstream in(path, std::ios::in | std::ios::binary);
  if (!in) {
    set_err(err, "cannot open " + path);
    return kErrOpen;
  }
  CsrHeader h{};
  if (!in.read(reinterpret_cast<char*>(&h), sizeof(h))) {
    set_err(err, "short CSR header in " + path);
    return kErrFormat;
  }
  if ((h.flags & (kVertexT64 | kEdgeT64)) || ((h.flags & kWeighted) && (h.flags & kWeightT64))) {
    set_err(err, "binary CSR header does not match the int32 / float types");
    return kErrTypes;
  }
  if (h.num_verts < 0 || h.num_edges < 0 || h.num_verts >= INT32_MAX || h.num_edges >= INT32_MAX) {
    set_err(err, "binary CSR sizes out of range in " + path);
    return kErrFormat;
  }
  out = Csr{};
  out.n = h.num_verts;
  out.flags = h.flags;
  out.off.resize((size_t)h.num_verts + 1);
  out.ind.resize((size_t)h.num_edges);
  in.read(reinterpret_cast<char*>(out.off.data()), (std::streamsize)(out.off.size() * sizeof(int32_t)));
  in.read(reinterpret_cast<char*>(out.ind.data()), (std::streamsize)(out.ind.size() * sizeof(int32_t)));
  if (h.flags & kWeighted) {
    out.w.resize((size_t)h.num_edges);
    in.read(reinterpret_cast<char*>(out.w.data()), (std::streamsize)(out.w.size() * sizeof(float)));
  }
  if (!in) {
    set_err(err, "truncated binary CSR " + path);
    return kErrFormat;
  }
  if (out.off[0] != 0 || out.off.back() != h.num_edges) {
    set_err(err, "binary CSR offsets do not span the edges in " + path);
    return kErrFormat;
  }
  return kOk;
}

int write_csr(const std::string& path, const Csr& g, std::string* err) {
  std::ofstream os(path, std::ios::out | std::ios::trunc | std::ios::binary);
  if (!os) {
    set_err(err, "cannot open " + path);
    return kErrOpen;
  }
  const bool weighted = (g.flags & kWeighted) != 0;
  CsrHeader h{kCsrVersion, g.n, (int64_t)g.ind.size(), g.flags & ~(uint64_t)(kVertexT64 | kEdgeT64 | kWeightT64)};
  os.write(reinterpret_cast<const char*>(&h), sizeof(h));
  os.write(reinterpret_cast<const char*>(g.off.data()), (std::streamsize)(g.off.size() * sizeof(int32_t)));
  os.write(reinterpret_cast<const char*>(g.ind.data()), (std::streamsize)(g.ind.size() * sizeof(int32_t)));
  if (weighted) {
    std::vector<float> w = g.w;
    w.resize(g.ind.size(), 1.0f);
    os.write(reinterpret_cast<const char*>(w.data()), (std::streamsize)(w.size() * sizeof(float)));
  }
  return os ? kOk : kErrFormat;
}

Csr csr_from_coo(const Coo& coo, bool add_reverse) {
  const size_t e = coo.src.size();
  std::vector<int32_t> s, d;
  std::vector<float> w;
  s.reserve(add_reverse ? 2 * e : e);
  d.reserve(s.capacity());
  w.reserve(s.capacity());
  int64_t n = coo.n;
  for (size_t k = 0; k < e; ++k) {
    n = std::max<int64_t>(n, (int64_t)std::max(coo.src[k], coo.dst[k]) + 1);
    if (coo.src[k] == coo.dst[k]) continue;  // no self loops
    s.push_back(coo.src[k]);
    d.push_back(coo.dst[k]);
    w.push_back(coo.w.empty() ? 1.0f : coo.w[k]);
    if (add_reverse) {
      s.push_back(coo.dst[k]);
      d.push_back(coo.src[k]);
      w.push_back(coo.w.empty() ? 1.0f : coo.w[k]);
    }
  }
  // stable order by (src, dst): the first occurrence of a repeated pair wins
  std::vector<size_t> ord(s.size());
  std::iota(ord.begin(), ord.end(), (size_t)0);
  std::stable_sort(ord.begin(), ord.end(),
                   [&](size_t x, size_t y) { return std::tie(s[x], d[x]) < std::tie(s[y], d[y]); });
  Csr g;
  g.n = n;
  g.off.assign((size_t)n + 1, 0);
  g.flags = (coo.weighted ? (uint64_t)kWeighted : 0) | kZeroIndexed | (coo.directed ? (uint64_t)kDirected : 0) |
            (add_reverse ? (uint64_t)kHasReverseEdges : 0);
  g.ind.reserve(ord.size());
  if (coo.weighted) g.w.reserve(ord.size());
  int64_t prev_s = -1, prev_d = -1;
  for (size_t k : ord) {
    if (s[k] == prev_s && d[k] == prev_d) continue;
    prev_s = s[k];
    prev_d = d[k];
    g.off[(size_t)s[k] + 1]++;
    g.ind.push_back(d[k]);
    if (coo.weighted) g.w.push_back(w[k]);
  }
  for (int64_t r = 0; r < n; ++r) g.off[(size_t)r + 1] += g.off[(size_t)r];
  return g;
}

Coo coo_from_csr(const Csr& g) {
  Coo c;
  c.n = g.n;
  c.weighted = (g.flags & kWeighted) != 0;
  c.directed = (g.flags & kDirected) != 0;
  c.src.reserve(g.ind.size());
  c.dst = g.ind;
  c.w.reserve(g.ind.size());
  for (int64_t r = 0; r < g.n; ++r)
    for (int32_t k = g.off[(size_t)r]; k < g.off[(size_t)r + 1]; ++k) {
      c.src.push_back((int32_t)r);
      c.w.push_back(g.w.empty() ? 1.0f : g.w[(size_t)k]);
    }
  return c;
}

void remove_reverse(Coo& coo) {
  // an entry (s, d, w) with s < d goes when (d, s, w) is present; entries
  // with s > d stay; self loops are their own reverse and go
  struct Key {
    int32_t a, b;
    uint32_t w;
    bool operator==(const Key& o) const { return a == o.a && b == o.b && w == o.w; }
  };
  struct KeyHash {
    size_t operator()(const Key& k) const {
      uint64_t h = ((uint64_t)(uint32_t)k.a << 32) ^ (uint32_t)k.b;
      h ^= (uint64_t)k.w * 0x9E3779B97F4A7C15ull;
      h ^= h >> 29;
      return (size_t)(h * 0xBF58476D1CE4E5B9ull);
    }
  };
  auto bits = [](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
  };
  std::unordered_set<Key, KeyHash> upper;
  upper.reserve(coo.src.size());
  for (size_t k = 0; k < coo.src.size(); ++k)
    if (coo.src[k] > coo.dst[k]) upper.insert(Key{coo.src[k], coo.dst[k], bits(coo.w[k])});
  size_t o = 0;
  for (size_t k = 0; k < coo.src.size(); ++k) {
    const bool keep = coo.src[k] > coo.dst[k] ||
                      (coo.src[k] < coo.dst[k] && !upper.count(Key{coo.dst[k], coo.src[k], bits(coo.w[k])}));
    if (!keep) continue;
    coo.src[o] = coo.src[k];
    coo.dst[o] = coo.dst[k];
    coo.w[o] = coo.w[k];
    ++o;
  }
  coo.src.resize(o);
  coo.dst.resize(o);
  coo.w.resize(o);
}

}  // namespace jcio

// C entry points for scripting and tests (host memory; no device involved).
extern "C" {

struct jcio_graph {
  jcio::Csr g;
};

// Reads .mtx or .csr into a CSR as the driver prepares it: undirected inputs
// without reverse edges get them; mtx_drop_weights mirrors the driver, which
// reads MTX weights as 1.0. Returns a handle or NULL (*status set).
jcio_graph* jcio_load(const char* path, int add_reverse_for_undirected, int mtx_drop_weights, int* status) {
  std::string err;
  int st = jcio::kOk;
  auto* h = new jcio_graph();
  const int kind = jcio::file_kind(path);
  if (kind == 0) {
    jcio::Coo c;
    st = jcio::read_mtx(path, c, mtx_drop_weights != 0, &err);
    if (st == jcio::kOk) h->g = jcio::csr_from_coo(c, add_reverse_for_undirected && !c.directed);
  } else if (kind == 1) {
    st = jcio::read_csr(path, h->g, &err);
    if (st == jcio::kOk && add_reverse_for_undirected && !(h->g.flags & jcio::kDirected) &&
        !(h->g.flags & jcio::kHasReverseEdges)) {
      jcio::Coo c = jcio::coo_from_csr(h->g);
      h->g = jcio::csr_from_coo(c, true);
    }
  } else {
    st = 1;
  }
  if (status) *status = st;
  if (st != jcio::kOk) {
    delete h;
    return nullptr;
  }
  return h;
}

int64_t jcio_n(const jcio_graph* h) { return h->g.n; }
int64_t jcio_m(const jcio_graph* h) { return (int64_t)h->g.ind.size(); }
uint64_t jcio_flags(const jcio_graph* h) { return h->g.flags; }
const int32_t* jcio_offsets(const jcio_graph* h) { return h->g.off.data(); }
const int32_t* jcio_indices(const jcio_graph* h) { return h->g.ind.data(); }
const float* jcio_weights(const jcio_graph* h) { return h->g.w.empty() ? nullptr : h->g.w.data(); }
void jcio_free(jcio_graph* h) { delete h; }

// Writes CSR arrays as .csr (flags as given) or .mtx (values = weights when
// given; remove_reverse keeps one direction per undirected pair).
int jcio_save(const char* path, int64_t n, int64_t m, const int32_t* off, const int32_t* ind, const float* w,
              uint64_t flags, int remove_reverse) {
  jcio::Csr g;
  g.n = n;
  g.off.assign(off, off + n + 1);
  g.ind.assign(ind, ind + m);
  if (w) g.w.assign(w, w + m);
  g.flags = flags | (w ? (uint64_t)jcio::kWeighted : 0);
  const int kind = jcio::file_kind(path);
  std::string err;
  if (kind == 1) {
    if (!remove_reverse) return jcio::write_csr(path, g, &err);
    jcio::Coo c = jcio::coo_from_csr(g);
    jcio::remove_reverse(c);
    jcio::Csr h = jcio::csr_from_coo(c, false);
    h.flags = g.flags & ~(uint64_t)jcio::kHasReverseEdges;
    return jcio::write_csr(path, h, &err);
  }
  if (kind == 0) {
    jcio::Coo c = jcio::coo_from_csr(g);
    if (remove_reverse) jcio::remove_reverse(c);
    return jcio::write_mtx(path, c, w != nullptr, (flags & jcio::kDirected) != 0, &err);
  }
  return 2;
}

}  // extern "C"
