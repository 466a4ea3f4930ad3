// Serving runtime: graphs resident in HBM as one arena, timed request replay,
// continuous batching over HIP streams under an HBM budget.
//
// No reference counterpart (the reference computes one graph per call,
// jaccard.cpp:926-955). Arithmetic is that of jc_jaccard / jc_jaccard_list:
// the owner engine accumulates exactly and every per-row reduction (row sums,
// per-pair waves) depends only on the row, so a request's bits do not depend
// on the batch it ran in.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "../../include/jaccard_amd.h"
#include "jc_device.hpp"
#include "jc_internal.hpp"

namespace jc {
hipError_t launch_symcheck(int32_t n, int64_t m, const int32_t* off, const int32_t* ind, int* flag,
                           hipStream_t s);
namespace {

// Arena builders: member graph g lands at rows [vbase, vbase + n), stored
// edges [ebase, ebase + m), column ids shifted by vbase.
__global__ void k_arena_offsets(int32_t n1, const int32_t* __restrict__ src, int32_t ebase,
                                int32_t* __restrict__ dst) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n1; i += gridDim.x * blockDim.x)
    dst[i] = src[i] + ebase;
}

__global__ void k_arena_indices(int64_t m, const int32_t* __restrict__ src, int32_t vbase,
                                int32_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i] + vbase;
}

// Result delivery: descriptor blockIdx.y copies count floats src -> dst.
struct ScatterDesc {
  const float* src;
  float* dst;
  int64_t count;
};

__global__ void k_scatter(const ScatterDesc* __restrict__ d) {
  const ScatterDesc e = d[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e.count; i += (int64_t)gridDim.x * blockDim.x)
    e.dst[i] = e.src[i];
}

size_t al(size_t v) { return (v + 255) / 256 * 256; }

// First-fit allocator over the budget region (offsets, 256-byte granules).
class Pool {
 public:
  void init(size_t cap) {
    cap_ = cap;
    used_ = peak_ = 0;
    free_.clear();
    if (cap) free_[0] = cap;
  }
  size_t largest() const {
    size_t b = 0;
    for (const auto& f : free_) b = std::max(b, f.second);
    return b;
  }
  bool alloc(size_t bytes, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t cap() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  void reset_peak() { peak_ = used_; }

 private:
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

struct Member {
  int32_t vbase, n;
  int64_t ebase, m;
  jc_csr local;
  const float* W;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  unsigned char* host = nullptr;  // pinned staging
  size_t host_bytes = 0;
  bool busy = false;
  size_t off = 0, bytes = 0;      // budget chunk
  int batch = -1;
  std::vector<int> reqs;
  std::vector<int> graphs;
};

// Byte layout of one batch (device chunk; the first four parts are staged
// from pinned host memory in one copy).
struct BatchLayout {
  size_t segs, desc, first, second, pout, ws, staged, total;
};

BatchLayout batch_layout(int nseg, int nd, int64_t K, int32_t R, int64_t E) {
  BatchLayout L{};
  size_t o = 0;
  L.segs = o;
  o += al(sizeof(int2) * (size_t)(nseg + 1));
  L.desc = o;
  o += al(sizeof(ScatterDesc) * (size_t)std::max(nd, 1));
  L.first = o;
  o += al(sizeof(int32_t) * (size_t)K);
  L.second = o;
  o += al(sizeof(int32_t) * (size_t)K);
  L.staged = o;
  L.pout = o;
  o += al(sizeof(float) * (size_t)K);
  L.ws = o;
  if (nseg > 0) o += al(engine_workspace_bytes(R, E));
  L.total = o;
  return L;
}

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void wait_until(std::chrono::steady_clock::time_point t0, double t) {
  for (;;) {
    const double d = t - now_s(t0);
    if (d <= 0) return;
    if (d > 2e-4)
      std::this_thread::sleep_for(std::chrono::duration<double>(d - 1e-4));
    else
      std::this_thread::yield();
  }
}

}  // namespace
}  // namespace jc

struct jc_server {
  int device = 0;
  bool weighted = false;
  int32_t N = 0;
  int64_t M = 0;
  int32_t* off = nullptr;
  int32_t* ind = nullptr;
  float* W = nullptr;
  double* S64 = nullptr;
  jc::ExactInfo* exact = nullptr;
  float* out = nullptr;
  size_t arena_bytes = 0;
  std::vector<jc::Member> g;
  jc_server_config cfg{};
  unsigned char* pool_mem = nullptr;
  jc::Pool pool;
  std::vector<jc::Slot> slots;
  hipStream_t fifo_stream = nullptr;
};

namespace {

using jc::note_hip_error;

#define SV_TRY(expr)                                   \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return note_hip_error(e_);   \
  } while (0)

void free_server(jc_server* s) {
  if (!s) return;
  for (auto& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.host) (void)hipHostFree(sl.host);
  }
  if (s->fifo_stream) {
    (void)hipStreamSynchronize(s->fifo_stream);
    (void)hipStreamDestroy(s->fifo_stream);
  }
  for (void* p : {(void*)s->off, (void*)s->ind, (void*)s->W, (void*)s->S64, (void*)s->exact, (void*)s->out,
                  (void*)s->pool_mem})
    if (p) (void)hipFree(p);
  delete s;
}

bool request_ok(const jc_server* s, const jc_request& q) {
  if (q.graph < 0 || q.graph >= (int32_t)s->g.size()) return false;
  if (q.kind == JC_REQ_FULL) return s->g[q.graph].m == 0 || q.result != nullptr;
  if (q.kind != JC_REQ_PAIRS || q.num_pairs < 0) return false;
  return q.num_pairs == 0 || (q.first && q.second && q.result);
}

// One request at a time with the plain entry points on the member's own CSR.
int run_fifo(jc_server* s, const jc_request* reqs, int32_t count, jc_request_stats* st,
             const std::vector<int>& order, std::chrono::steady_clock::time_point t0) {
  hipStream_t hs = s->fifo_stream;
  jc_stream_t js = reinterpret_cast<jc_stream_t>(hs);
  int b = 0;
  for (int i : order) {
    const jc_request& q = reqs[i];
    jc_request_stats& r = st[i];
    jc::wait_until(t0, q.arrival);
    r.admitted = jc::now_s(t0);
    r.batch = b++;
    if (!request_ok(s, q)) {
      r.status = JC_ERR_INVALID_ARG;
      r.completed = r.admitted;
      continue;
    }
    const jc::Member& m = s->g[q.graph];
    int code = JC_OK;
    if (q.kind == JC_REQ_FULL) {
      code = jc_jaccard(&m.local, m.W, q.result, JC_EDGE_CENTRIC, js);
    } else if (q.num_pairs > 0) {
      const size_t k = (size_t)q.num_pairs;
      int32_t* d = nullptr;
      hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&d), 2 * k * sizeof(int32_t), hs);
      if (e == hipSuccess) e = hipMemcpyAsync(d, q.first, k * sizeof(int32_t), hipMemcpyHostToDevice, hs);
      if (e == hipSuccess) e = hipMemcpyAsync(d + k, q.second, k * sizeof(int32_t), hipMemcpyHostToDevice, hs);
      if (e == hipSuccess) code = jc_jaccard_list(&m.local, m.W, q.num_pairs, d, d + k, q.result, js);
      else code = note_hip_error(e);
      if (d) (void)hipFreeAsync(d, hs);
    }
    const hipError_t e = hipStreamSynchronize(hs);
    if (code == JC_OK && e != hipSuccess) code = note_hip_error(e);
    r.status = code;
    r.completed = jc::now_s(t0);
  }
  return JC_OK;
}

}  // namespace

extern "C" {

int jc_server_create(int32_t num_graphs, const jc_csr* graphs, const float* const* vertex_weights,
                     const jc_server_config* cfg_in, jc_server** out) {
  if (!out || num_graphs < 1 || !graphs) return JC_ERR_INVALID_ARG;
  *out = nullptr;
  jc_server_config cfg = cfg_in ? *cfg_in : jc_server_config{};
  if (cfg.slots == 0) cfg.slots = 4;
  if (cfg.max_batch_requests == 0) cfg.max_batch_requests = 256;
  if (cfg.max_batch_edges == 0) cfg.max_batch_edges = int64_t(1) << 26;
  if (cfg.max_batch_pairs == 0) cfg.max_batch_pairs = int64_t(1) << 20;
  if (cfg.hbm_budget_bytes == 0) cfg.hbm_budget_bytes = size_t(256) << 20;
  if (cfg.slots < 1 || cfg.slots > 16 || cfg.max_batch_requests < 1 || cfg.max_batch_requests > 65535 ||
      cfg.max_batch_edges < 1 || cfg.max_batch_pairs < 1 ||
      (cfg.mode != JC_SERVE_CONTINUOUS && cfg.mode != JC_SERVE_FIFO))
    return JC_ERR_INVALID_ARG;

  jc_server* s = new jc_server();
  s->cfg = cfg;
  s->weighted = vertex_weights != nullptr;
  int64_t N = 0, M = 0;
  for (int32_t i = 0; i < num_graphs; ++i) {
    const jc_csr& c = graphs[i];
    if (c.n < 0 || c.m < 0 || (c.n > 0 && !c.offsets) || (c.m > 0 && !c.indices) ||
        (s->weighted && c.n > 0 && !vertex_weights[i])) {
      delete s;
      return JC_ERR_INVALID_ARG;
    }
    jc::Member m{(int32_t)N, c.n, M, c.m, c, s->weighted ? vertex_weights[i] : nullptr};
    s->g.push_back(m);
    N += c.n;
    M += c.m;
  }
  if (N >= INT32_MAX || M >= INT32_MAX) {  // int32 CSR arena
    delete s;
    return JC_ERR_INVALID_ARG;
  }
  s->N = (int32_t)N;
  s->M = M;
  auto fail = [&](hipError_t e) {
    const int code = note_hip_error(e);
    free_server(s);
    return code;
  };
  hipError_t e;
  if ((e = hipGetDevice(&s->device)) != hipSuccess) return fail(e);
  if ((e = hipStreamCreateWithFlags(&s->fifo_stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
  const size_t nb = (size_t)(N + 1) * sizeof(int32_t), mb = (size_t)std::max<int64_t>(M, 1) * sizeof(int32_t);
  if ((e = hipMalloc(&s->off, nb)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&s->ind, mb)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&s->out, mb)) != hipSuccess) return fail(e);
  s->arena_bytes = nb + 2 * mb;
  if (s->weighted) {
    if ((e = hipMalloc(&s->W, (size_t)std::max<int64_t>(N, 1) * sizeof(float))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&s->S64, (size_t)std::max<int64_t>(N, 1) * sizeof(double))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&s->exact, sizeof(jc::ExactInfo))) != hipSuccess) return fail(e);
    s->arena_bytes += (size_t)std::max<int64_t>(N, 1) * (sizeof(float) + sizeof(double));
  }
  hipStream_t st = s->fifo_stream;
  if ((e = hipMemsetAsync(s->off, 0, sizeof(int32_t), st)) != hipSuccess) return fail(e);
  for (const jc::Member& m : s->g) {
    if (m.n > 0) {
      const int gb = (int)std::min<int64_t>((m.n + 1 + 255) / 256, 4096);
      hipLaunchKernelGGL(jc::k_arena_offsets, dim3(gb), dim3(256), 0, st, m.n + 1, m.local.offsets,
                         (int32_t)m.ebase, s->off + m.vbase);
      if ((e = hipGetLastError()) != hipSuccess) return fail(e);
      if (s->weighted &&
          (e = hipMemcpyAsync(s->W + m.vbase, m.W, (size_t)m.n * sizeof(float), hipMemcpyDeviceToDevice, st)) !=
              hipSuccess)
        return fail(e);
    }
    if (m.m > 0) {
      const int gb = (int)std::min<int64_t>((m.m + 255) / 256, 8192);
      hipLaunchKernelGGL(jc::k_arena_indices, dim3(gb), dim3(256), 0, st, m.m, m.local.indices, m.vbase,
                         s->ind + m.ebase);
      if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    }
  }
  // The batched engine pass runs without the fix-up of unmatched positions,
  // so every member must be symmetric: the arena (their disjoint union) is
  // checked once here.
  if (M > 0) {
    int* flag = nullptr;
    if ((e = hipMalloc(&flag, sizeof(int))) != hipSuccess) return fail(e);
    int h = 0;
    e = jc::launch_fill_u32(reinterpret_cast<uint32_t*>(flag), 1, 0u, st);
    if (e == hipSuccess) e = jc::launch_symcheck(s->N, s->M, s->off, s->ind, flag, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const hipError_t ef = hipFree(flag);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess) return fail(e);
    if (h != 0) {
      free_server(s);
      return JC_ERR_INVALID_CSR;
    }
  }
  if (s->weighted && N > 0 &&
      (e = jc::launch_rowsum_f64(s->N, s->off, s->ind, s->W, s->S64, nullptr, st)) != hipSuccess)
    return fail(e);
  if (s->weighted && (e = jc::launch_exact_info(s->N, s->W, s->S64, s->exact, st)) != hipSuccess) return fail(e);
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e);

  if (cfg.mode == JC_SERVE_CONTINUOUS) {
    if ((e = hipMalloc(&s->pool_mem, cfg.hbm_budget_bytes)) != hipSuccess) return fail(e);
    s->pool.init(cfg.hbm_budget_bytes / 256 * 256);
    s->slots.resize(cfg.slots);
    for (auto& sl : s->slots) {
      if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
      if ((e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess) return fail(e);
    }
  }
  *out = s;
  return JC_OK;
}

int jc_server_destroy(jc_server* s) {
  if (!s) return JC_ERR_INVALID_ARG;
  free_server(s);
  return JC_OK;
}

size_t jc_server_arena_bytes(const jc_server* s) { return s ? s->arena_bytes : 0; }

int jc_server_run(jc_server* s, const jc_request* reqs, int32_t count, jc_request_stats* stats,
                  jc_server_run_info* info) {
  if (!s || count < 0 || (count > 0 && (!reqs || !stats))) return JC_ERR_INVALID_ARG;
  SV_TRY(hipSetDevice(s->device));
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) {
    order[i] = i;
    stats[i] = jc_request_stats{0.0, 0.0, -1, JC_OK};
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival < reqs[b].arrival; });
  jc_server_run_info inf{};
  const auto t0 = std::chrono::steady_clock::now();

  if (s->cfg.mode == JC_SERVE_FIFO) {
    const int code = run_fifo(s, reqs, count, stats, order, t0);
    inf.batches = count;
    inf.max_in_flight = count > 0 ? 1 : 0;
    inf.wall_seconds = jc::now_s(t0);
    if (info) *info = inf;
    return code;
  }

  const jc_server_config& cfg = s->cfg;
  const int mode = s->weighted ? jc::kModeEcW : jc::kModeEcU;
  s->pool.reset_peak();
  std::vector<char> graph_busy(s->g.size(), 0), in_batch(s->g.size(), 0);
  std::deque<int> pending;
  int next = 0, finished = 0, batches = 0, in_flight = 0;
  int err = JC_OK;

  auto finish = [&](int i, int status) {
    stats[i].completed = jc::now_s(t0);
    stats[i].status = status;
    ++finished;
  };

  // Forms and launches one batch on slot sl from the pending queue; false if
  // nothing could be admitted.
  auto launch = [&](jc::Slot& sl) -> bool {
    std::vector<int> take;
    int nseg = 0;
    int32_t R = 0;
    int64_t E = 0, K = 0;
    for (auto it = pending.begin(); it != pending.end() && (int)take.size() < cfg.max_batch_requests;) {
      const jc_request& q = reqs[*it];
      const jc::Member& m = s->g[q.graph];
      const bool full = q.kind == JC_REQ_FULL;
      if (full && (graph_busy[q.graph] || in_batch[q.graph])) { ++it; continue; }
      if (!take.empty() && full && E + m.m > cfg.max_batch_edges) { ++it; continue; }
      if (!take.empty() && !full && K + q.num_pairs > cfg.max_batch_pairs) { ++it; continue; }
      const jc::BatchLayout L = jc::batch_layout(nseg + (full ? 1 : 0), (int)take.size() + 1, K + (full ? 0 : q.num_pairs),
                                                 R + (full ? m.n : 0), E + (full ? m.m : 0));
      if (L.total > s->pool.largest()) {
        if (take.empty() && s->pool.used() == 0) {  // can never fit the budget
          finish(*it, JC_ERR_OUT_OF_MEMORY);
          stats[*it].admitted = stats[*it].completed;
          it = pending.erase(it);
          continue;
        }
        break;  // wait for memory, in arrival order
      }
      if (full) {
        in_batch[q.graph] = 1;
        ++nseg;
        R += m.n;
        E += m.m;
      } else {
        K += q.num_pairs;
      }
      take.push_back(*it);
      it = pending.erase(it);
    }
    if (take.empty()) return false;
    const int nd = (int)take.size();
    const jc::BatchLayout L = jc::batch_layout(nseg, nd, K, R, E);
    size_t off = 0;
    if (!s->pool.alloc(L.total, &off)) {  // cannot happen after the largest() checks
      for (int i : take) {
        if (reqs[i].kind == JC_REQ_FULL) in_batch[reqs[i].graph] = 0;
      }
      pending.insert(pending.begin(), take.begin(), take.end());
      return false;
    }
    unsigned char* dev = s->pool_mem + off;
    if (sl.host_bytes < L.staged) {
      const size_t want = std::max(L.staged, sl.host_bytes * 2);
      if (sl.host) (void)hipHostFree(sl.host);
      sl.host = nullptr;
      sl.host_bytes = 0;
      hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&sl.host), want, hipHostMallocDefault);
      if (e != hipSuccess) {
        err = note_hip_error(e);
        s->pool.release(off, L.total);
        for (int i : take) {
          if (reqs[i].kind == JC_REQ_FULL) in_batch[reqs[i].graph] = 0;
          stats[i].admitted = jc::now_s(t0);
          finish(i, err);
        }
        return true;
      }
      sl.host_bytes = want;
    }
    int2* hsegs = reinterpret_cast<int2*>(sl.host + L.segs);
    jc::ScatterDesc* hdesc = reinterpret_cast<jc::ScatterDesc*>(sl.host + L.desc);
    int32_t* hfirst = reinterpret_cast<int32_t*>(sl.host + L.first);
    int32_t* hsecond = reinterpret_cast<int32_t*>(sl.host + L.second);
    const float* dpout = reinterpret_cast<const float*>(dev + L.pout);
    int si = 0;
    int32_t rows = 0;
    int64_t k0 = 0, maxcount = 0;
    sl.graphs.clear();
    for (int d = 0; d < nd; ++d) {
      const jc_request& q = reqs[take[d]];
      const jc::Member& m = s->g[q.graph];
      if (q.kind == JC_REQ_FULL) {
        hsegs[si++] = make_int2(m.vbase, rows);
        rows += m.n;
        hdesc[d] = jc::ScatterDesc{s->out + m.ebase, q.result, m.m};
        maxcount = std::max(maxcount, m.m);
        graph_busy[q.graph] = 1;
        in_batch[q.graph] = 0;
        sl.graphs.push_back(q.graph);
      } else {
        for (int32_t p = 0; p < q.num_pairs; ++p) {
          const int32_t a = q.first[p], b = q.second[p];
          hfirst[k0 + p] = (a >= 0 && a < m.n) ? a + m.vbase : -1;
          hsecond[k0 + p] = (b >= 0 && b < m.n) ? b + m.vbase : -1;
        }
        hdesc[d] = jc::ScatterDesc{dpout + k0, q.result, q.num_pairs};
        maxcount = std::max<int64_t>(maxcount, q.num_pairs);
        k0 += q.num_pairs;
      }
    }
    hsegs[nseg] = make_int2(0, rows);
    const double tadm = jc::now_s(t0);
    for (int i : take) {
      stats[i].admitted = tadm;
      stats[i].batch = batches;
    }
    hipStream_t st = sl.stream;
    hipError_t e = hipMemcpyAsync(dev, sl.host, L.staged, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nseg > 0) {
      jc::EngineArgs a{};
      a.n = s->N;
      a.m = s->M;
      a.off = s->off;
      a.ind = s->ind;
      a.vw = s->W;
      a.S64 = s->S64;
      a.exact = s->exact;
      a.out32 = s->out;
      a.segs = reinterpret_cast<const int2*>(dev + L.segs);
      a.nseg = nseg;
      a.seg_rows = R;
      a.seg_edges = E;
      e = jc::run_engine(mode, s->weighted, a, dev + L.ws, false, st);
    }
    if (e == hipSuccess && K > 0)
      e = jc::run_pairs(mode, s->weighted, s->N, s->off, s->ind, s->W, nullptr, K,
                        reinterpret_cast<const int32_t*>(dev + L.first), reinterpret_cast<const int32_t*>(dev + L.second),
                        reinterpret_cast<float*>(dev + L.pout), nullptr, 0, st);
    if (e == hipSuccess && maxcount > 0) {
      const int gx = (int)std::max<int64_t>(1, std::min<int64_t>((maxcount + 1023) / 1024, 2048));
      hipLaunchKernelGGL(jc::k_scatter, dim3(gx, nd), dim3(256), 0, st,
                         reinterpret_cast<const jc::ScatterDesc*>(dev + L.desc));
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(sl.done, st);
    sl.busy = true;
    sl.off = off;
    sl.bytes = L.total;
    sl.batch = batches++;
    sl.reqs = take;
    if (e != hipSuccess) err = note_hip_error(e);  // the loop below retires everything
    return true;
  };

  while (finished < count) {
    if (err != JC_OK) {  // a HIP call failed: drain the device, fail what is left
      for (auto& sl : s->slots) {
        if (!sl.busy) continue;
        (void)hipStreamSynchronize(sl.stream);
        for (int i : sl.reqs) finish(i, err);
        s->pool.release(sl.off, sl.bytes);
        sl.busy = false;
      }
      for (int i : pending) finish(i, err);
      pending.clear();
      for (; next < count; ++next) {
        stats[order[next]].admitted = jc::now_s(t0);
        finish(order[next], err);
      }
      break;
    }
    bool progress = false;
    const double t = jc::now_s(t0);
    while (next < count && reqs[order[next]].arrival <= t) {
      const int i = order[next++];
      const jc_request& q = reqs[i];
      if (!request_ok(s, q) || (q.kind == JC_REQ_PAIRS && q.num_pairs == 0) ||
          (q.kind == JC_REQ_FULL && s->g[q.graph].m == 0)) {
        stats[i].admitted = t;
        finish(i, request_ok(s, q) ? JC_OK : JC_ERR_INVALID_ARG);
      } else {
        pending.push_back(i);
      }
      progress = true;
    }
    int busy = 0;
    for (auto& sl : s->slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) {
        ++busy;
        continue;
      }
      if (q != hipSuccess) {
        err = note_hip_error(q);
        ++busy;
        break;
      }
      const int status = JC_OK;
      for (int i : sl.reqs) finish(i, status);
      for (int gi : sl.graphs) graph_busy[gi] = 0;
      s->pool.release(sl.off, sl.bytes);
      sl.busy = false;
      progress = true;
    }
    for (auto& sl : s->slots) {
      if (pending.empty()) break;
      if (sl.busy) continue;
      if (!launch(sl)) break;
      if (sl.busy) ++busy;
      progress = true;
    }
    in_flight = std::max(in_flight, busy);
    if (!progress) {
      if (busy == 0 && pending.empty() && next < count)
        jc::wait_until(t0, reqs[order[next]].arrival);
      else
        std::this_thread::yield();
    }
  }
  inf.batches = batches;
  inf.max_in_flight = in_flight;
  inf.peak_budget_bytes = s->pool.peak();
  inf.wall_seconds = jc::now_s(t0);
  if (info) *info = inf;
  return err;
}

}  // extern "C"
