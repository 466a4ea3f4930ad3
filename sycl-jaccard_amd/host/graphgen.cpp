// Deterministic synthetic graph generators and the CSR builder used by the
// tests, the benchmark, the batch packer and the serving registry.
//
// Reference semantics for the CSR the Jaccard kernels consume:
//   * main.cpp:98-120           undirected input is symmetrised (reverse edges added)
//   * readMtxToCSR.cpp:152-208  std::set dedupes (u,v), self loops are dropped,
//                               columns are sorted inside each row
// The builder here produces the same CSR without std::set: a parallel counting
// sort by source row, then a per-row sort + unique. Unlike mtxSetToCSR it keeps
// every vertex id in [0,n) as a row (isolated vertices get empty rows).
//
// Randomness is counter based (SplitMix64 finaliser of (seed, edge, draw)), so
// a graph is a pure function of its parameters: identical on every machine and
// for every thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>
#include <omp.h>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t idx) {
  return mix64(mix64(seed ^ 0x5851F42D4C957F2Dull) ^ mix64(stream * 0xD1B54A32D192ED03ull + idx));
}

// Symmetric per-pair weight in (0,1], exactly representable in fp32:
// ((h >> 40) + 1) * 2^-24 with h a hash of (seed, min(u,v), max(u,v)).
inline float pair_weight(uint64_t seed, int32_t u, int32_t v) {
  uint32_t lo = (uint32_t)std::min(u, v), hi = (uint32_t)std::max(u, v);
  uint64_t h = draw(seed, 0x77656967687473ull, ((uint64_t)lo << 32) | hi);
  return (float)((double)((h >> 40) + 1) * (1.0 / 16777216.0));
}

}  // namespace

struct jcg_graph {
  int32_t n = 0;
  int64_t m = 0;
  std::vector<int32_t> offsets;
  std::vector<int32_t> indices;
  std::vector<float> weights;  // empty when unweighted
  bool weighted = false;
};

namespace {

// Counting sort by source row + per-row sort/unique. src/dst hold e directed
// edges; symmetrize adds the reverse of each; self loops are always dropped.
jcg_graph* build_csr(int32_t n, int64_t e, const int32_t* src, const int32_t* dst, bool symmetrize) {
  std::vector<int64_t> cnt((size_t)n + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < e; ++i) {
    int32_t u = src[i], v = dst[i];
    if (u == v) continue;
    __atomic_fetch_add(&cnt[u], 1, __ATOMIC_RELAXED);
    if (symmetrize) __atomic_fetch_add(&cnt[v], 1, __ATOMIC_RELAXED);
  }
  std::vector<int64_t> start((size_t)n + 1, 0);
  for (int32_t r = 0; r < n; ++r) start[r + 1] = start[r] + cnt[r];
  int64_t total = start[n];
  std::vector<int32_t> tmp((size_t)total);
  std::vector<int64_t> cur(start.begin(), start.end() - 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < e; ++i) {
    int32_t u = src[i], v = dst[i];
    if (u == v) continue;
    tmp[__atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED)] = v;
    if (symmetrize) tmp[__atomic_fetch_add(&cur[v], 1, __ATOMIC_RELAXED)] = u;
  }
  std::vector<int64_t> deg((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t r = 0; r < n; ++r) {
    int32_t* b = tmp.data() + start[r];
    int32_t* en = tmp.data() + start[r + 1];
    std::sort(b, en);
    deg[r] = std::unique(b, en) - b;
  }
  auto* g = new jcg_graph;
  g->n = n;
  g->offsets.assign((size_t)n + 1, 0);
  int64_t acc = 0;
  for (int32_t r = 0; r < n; ++r) {
    if (acc > INT32_MAX) break;
    g->offsets[r] = (int32_t)acc;
    acc += deg[r];
  }
  if (acc > INT32_MAX) {  // int32 CSR cannot hold this graph
    delete g;
    return nullptr;
  }
  g->offsets[n] = (int32_t)acc;
  g->m = acc;
  g->indices.resize((size_t)acc);
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t r = 0; r < n; ++r)
    std::memcpy(g->indices.data() + g->offsets[r], tmp.data() + start[r], sizeof(int32_t) * deg[r]);
  return g;
}

void attach_weights(jcg_graph* g, uint64_t wseed) {
  g->weighted = true;
  g->weights.resize((size_t)g->m);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t r = 0; r < g->n; ++r)
    for (int32_t k = g->offsets[r]; k < g->offsets[r + 1]; ++k)
      g->weights[k] = pair_weight(wseed, r, g->indices[k]);
}

// One RMAT endpoint pair for edge index `eid`, vertex ids below n (n need not be
// a power of two: out-of-range draws are redrawn with the next attempt counter).
inline void rmat_edge(uint64_t seed, int scale, int32_t n, uint64_t eid, uint32_t ta, uint32_t tab,
                      uint32_t tabc, int32_t* pu, int32_t* pv) {
  for (uint64_t attempt = 0;; ++attempt) {
    uint32_t u = 0, v = 0;
    uint64_t stream = eid * 64 + attempt;
    uint64_t bits = 0;
    for (int l = 0; l < scale; ++l) {
      if ((l & 1) == 0) bits = draw(seed, stream, (uint64_t)(l >> 1));
      uint32_t r = (uint32_t)(bits >> ((l & 1) ? 32 : 0));
      uint32_t bu = r >= tab ? 1u : 0u;  // quadrants c, d put the edge in the lower half
      uint32_t bv = (r >= ta && r < tab) || r >= tabc ? 1u : 0u;  // b, d: right half
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if ((int64_t)u < n && (int64_t)v < n) {
      *pu = (int32_t)u;
      *pv = (int32_t)v;
      return;
    }
  }
}

}  // namespace

extern "C" {

// RMAT (Chakrabarti et al. 2004) with quadrant probabilities a, b, c and
// d = 1-a-b-c; num_edges = round(edge_factor * n) draws; symmetrised, deduped,
// self loops dropped, columns sorted. weighted != 0 attaches symmetric per-pair
// weights in (0,1] drawn from wseed. n = 2^scale unless n_override > 0.
jcg_graph* jcg_rmat(int scale, int32_t n_override, double edge_factor, double a, double b, double c,
                    uint64_t seed, int weighted, uint64_t wseed) {
  if (scale < 1 || scale > 30) return nullptr;
  int32_t n = n_override > 0 ? n_override : (int32_t)(1u << scale);
  if ((int64_t)n > (1ll << scale)) return nullptr;
  int64_t e = (int64_t)std::llround(edge_factor * (double)n);
  const double two32 = 4294967296.0;
  uint32_t ta = (uint32_t)std::min(two32 - 1, std::floor(a * two32));
  uint32_t tab = (uint32_t)std::min(two32 - 1, std::floor((a + b) * two32));
  uint32_t tabc = (uint32_t)std::min(two32 - 1, std::floor((a + b + c) * two32));
  std::vector<int32_t> src((size_t)e), dst((size_t)e);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < e; ++i) rmat_edge(seed, scale, n, (uint64_t)i, ta, tab, tabc, &src[i], &dst[i]);
  jcg_graph* g = build_csr(n, e, src.data(), dst.data(), true);
  if (g && weighted) attach_weights(g, wseed);
  return g;
}

// The raw RMAT draws of jcg_rmat (same seed and parameters give the same
// edges, before symmetrisation and dedupe): src/dst caller-allocated with
// round(edge_factor * n) entries. Returns that count, or -1.
int64_t jcg_rmat_edges(int scale, int32_t n_override, double edge_factor, double a, double b, double c,
                       uint64_t seed, int32_t* src, int32_t* dst, int64_t capacity) {
  if (scale < 1 || scale > 30) return -1;
  int32_t n = n_override > 0 ? n_override : (int32_t)(1u << scale);
  if ((int64_t)n > (1ll << scale)) return -1;
  int64_t e = (int64_t)std::llround(edge_factor * (double)n);
  if (e > capacity) return -1;
  const double two32 = 4294967296.0;
  uint32_t ta = (uint32_t)std::min(two32 - 1, std::floor(a * two32));
  uint32_t tab = (uint32_t)std::min(two32 - 1, std::floor((a + b) * two32));
  uint32_t tabc = (uint32_t)std::min(two32 - 1, std::floor((a + b + c) * two32));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < e; ++i) rmat_edge(seed, scale, n, (uint64_t)i, ta, tab, tabc, &src[i], &dst[i]);
  return e;
}

// Chung-Lu graph with a power-law expected-degree sequence
//   w_i = w_max * (1 + i/i0)^(-1/(gamma-1)),  i in [0,n)
// where i0 is solved so that mean(w) = mean_deg. Endpoints of each sampled edge
// are drawn independently with P(i) = w_i / sum(w). Because dedupe removes
// repeated hub-hub pairs, the number of samples is calibrated (three secant
// steps, deterministic) so the final mean stored degree is close to mean_deg.
jcg_graph* jcg_chung_lu(int32_t n, double gamma, double mean_deg, double max_deg, uint64_t seed,
                        int weighted, uint64_t wseed) {
  if (n < 2 || gamma <= 1.0 || mean_deg <= 0 || max_deg < mean_deg) return nullptr;
  const double beta = 1.0 / (gamma - 1.0);
  auto mean_for = [&](double i0) {
    double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int32_t i = 0; i < n; ++i) s += max_deg * std::pow(1.0 + i / i0, -beta);
    return s / n;
  };
  double lo = 1e-6, hi = (double)n * 1e3;
  for (int it = 0; it < 64; ++it) {  // mean_for is increasing in i0
    double mid = std::sqrt(lo * hi);
    if (mean_for(mid) < mean_deg) lo = mid; else hi = mid;
  }
  const double i0 = std::sqrt(lo * hi);
  std::vector<double> cdf((size_t)n);
  double acc = 0;
  for (int32_t i = 0; i < n; ++i) {
    acc += max_deg * std::pow(1.0 + i / i0, -beta);
    cdf[i] = acc;
  }
  const double total = acc;
  auto pick = [&](uint64_t h) {
    double x = (double)(h >> 11) * (1.0 / 9007199254740992.0) * total;
    int32_t k = (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    return std::min(k, n - 1);
  };
  auto sample = [&](int64_t e) {
    std::vector<int32_t> src((size_t)e), dst((size_t)e);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < e; ++i) {
      src[i] = pick(draw(seed, 0x436c, 2 * (uint64_t)i));
      dst[i] = pick(draw(seed, 0x436c, 2 * (uint64_t)i + 1));
    }
    return build_csr(n, e, src.data(), dst.data(), true);
  };
  const double target = mean_deg * n;  // stored (directed) edges wanted
  int64_t e0 = (int64_t)(target / 2), e1 = 0;
  jcg_graph* g = sample(e0);
  if (!g) return nullptr;
  double m0 = (double)g->m;
  for (int it = 0; it < 3; ++it) {
    if (std::fabs(m0 - target) <= 0.002 * target) break;
    e1 = (int64_t)std::llround((double)e0 * target / m0);
    if (e1 == e0) break;
    delete g;
    g = sample(e1);
    if (!g) return nullptr;
    double m1 = (double)g->m;
    e0 = e1;
    m0 = m1;
  }
  if (weighted) attach_weights(g, wseed);
  return g;
}

// Build a CSR from an explicit directed edge list. symmetrize != 0 adds reverse
// edges (reference main.cpp:99-112); duplicates and self loops are dropped.
jcg_graph* jcg_from_edges(int32_t n, int64_t e, const int32_t* src, const int32_t* dst,
                          int symmetrize, int weighted, uint64_t wseed) {
  if (n < 0 || e < 0) return nullptr;
  for (int64_t i = 0; i < e; ++i)
    if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) return nullptr;
  jcg_graph* g = build_csr(n, e, src, dst, symmetrize != 0);
  if (g && weighted) attach_weights(g, wseed);
  return g;
}

int32_t jcg_n(const jcg_graph* g) { return g->n; }
int64_t jcg_m(const jcg_graph* g) { return g->m; }
int jcg_weighted(const jcg_graph* g) { return g->weighted ? 1 : 0; }
const int32_t* jcg_offsets(const jcg_graph* g) { return g->offsets.data(); }
const int32_t* jcg_indices(const jcg_graph* g) { return g->indices.data(); }
const float* jcg_weights(const jcg_graph* g) { return g->weights.empty() ? nullptr : g->weights.data(); }
void jcg_free(jcg_graph* g) { delete g; }

// A batch of RMAT graphs packed as one block-diagonal CSR (the disjoint
// union, global vertex ids): graph i has n[i] vertices, round(ef[i] * n[i])
// draws from seed seeds[i], edge weights when weighted != 0 (weight seed
// seeds[i] ^ wxor). vbase[i] / ebase[i] (count + 1 entries, caller-allocated)
// receive each graph's first vertex / first stored edge in the union.
// Graphs are generated in parallel, largest first. Returns nullptr if the
// union does not fit int32 (> 2^31 - 1 vertices or stored edges).
jcg_graph* jcg_rmat_batch(int32_t count, const int32_t* n, const double* ef, const uint64_t* seeds, int weighted,
                          uint64_t wxor, double a, double b, double c, int64_t* vbase, int64_t* ebase) {
  if (count <= 0) return nullptr;
  std::vector<jcg_graph*> parts((size_t)count, nullptr);
  std::vector<int32_t> order((size_t)count);
  for (int32_t i = 0; i < count; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return (double)n[x] * ef[x] > (double)n[y] * ef[y]; });
  const int nested = omp_get_max_active_levels();
  omp_set_max_active_levels(1);
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t t = 0; t < count; ++t) {
    const int32_t i = order[t];
    int scale = 1;
    while ((1ll << scale) < n[i]) ++scale;
    parts[i] = jcg_rmat(scale, n[i], ef[i], a, b, c, seeds[i], weighted, seeds[i] ^ wxor);
  }
  omp_set_max_active_levels(nested);
  bool ok = true;
  vbase[0] = 0;
  ebase[0] = 0;
  for (int32_t i = 0; i < count; ++i) {
    if (!parts[i]) ok = false;
    vbase[i + 1] = vbase[i] + (parts[i] ? parts[i]->n : 0);
    ebase[i + 1] = ebase[i] + (parts[i] ? parts[i]->m : 0);
  }
  if (!ok || vbase[count] > INT32_MAX || ebase[count] > INT32_MAX) {
    for (auto* p : parts) delete p;
    return nullptr;
  }
  auto* u = new jcg_graph;
  u->n = (int32_t)vbase[count];
  u->m = ebase[count];
  u->weighted = weighted != 0;
  u->offsets.resize((size_t)u->n + 1);
  u->indices.resize((size_t)u->m);
  if (u->weighted) u->weights.resize((size_t)u->m);
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t i = 0; i < count; ++i) {
    const jcg_graph* g = parts[i];
    const int32_t vb = (int32_t)vbase[i], eb = (int32_t)ebase[i];
    for (int32_t r = 0; r < g->n; ++r) u->offsets[vb + r] = eb + g->offsets[r];
    for (int64_t k = 0; k < g->m; ++k) u->indices[eb + k] = vb + g->indices[k];
    if (u->weighted) std::memcpy(u->weights.data() + eb, g->weights.data(), sizeof(float) * g->m);
  }
  u->offsets[u->n] = (int32_t)u->m;
  for (auto* p : parts) delete p;
  return u;
}

// The per-pair weight function, exported so tests can check symmetry/range.
float jcg_pair_weight(uint64_t wseed, int32_t u, int32_t v) { return pair_weight(wseed, u, v); }

// Exact int64 statistics used by the roofline figure:
//   out[0] = sum over stored edges (u,v) of min(d_u, d_v)
//   out[1] = sum over stored edges of (d_u + d_v)
//   out[2] = max degree, out[3] = number of isolated vertices
void jcg_stats(int32_t n, const int32_t* offsets, const int32_t* indices, int64_t* out) {
  int64_t smin = 0, ssum = 0, dmax = 0, iso = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : smin, ssum, iso) reduction(max : dmax)
  for (int32_t u = 0; u < n; ++u) {
    int64_t du = offsets[u + 1] - offsets[u];
    if (du == 0) ++iso;
    if (du > dmax) dmax = du;
    for (int32_t k = offsets[u]; k < offsets[u + 1]; ++k) {
      int32_t v = indices[k];
      int64_t dv = offsets[v + 1] - offsets[v];
      smin += std::min(du, dv);
      ssum += du + dv;
    }
  }
  out[0] = smin;
  out[1] = ssum;
  out[2] = dmax;
  out[3] = iso;
}

}  // extern "C"
