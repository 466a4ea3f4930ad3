// C ABI entry points (include/jaccard_amd.h) over the gfx950 kernels.
//
// Orchestration follows sygraph::detail::jaccard (jaccard.cpp:784-899):
//   edge-centric   : [row sums when weighted] -> owner engine, fused epilogue
//   vertex-centric : row sums -> owner engine writing weight_i/weight_s -> jw
// Temporaries come from a caller workspace or stream-ordered allocations.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "../../include/jaccard_amd.h"
#include "jc_device.hpp"
#include "jc_internal.hpp"

namespace jc {
hipError_t launch_validate_rows(int32_t n, const int32_t* off, const int32_t* ind, int* flags,
                                hipStream_t s);
hipError_t launch_symcheck(int32_t n, int64_t m, const int32_t* off, const int32_t* ind, int* flag,
                           hipStream_t s);
}  // namespace jc

namespace {

thread_local int g_last_hip = 0;

int hip_fail(hipError_t e) {
  g_last_hip = (int)e;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return JC_ERR_OUT_OF_MEMORY;
  return JC_ERR_HIP;
}

#define JC_TRY(expr)                          \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) return hip_fail(e_); \
  } while (0)

size_t al(size_t v) { return (v + 255) / 256 * 256; }

hipStream_t hs(jc_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

bool csr_args_ok(const jc_csr* g) {
  if (!g || g->n < 0 || g->m < 0) return false;
  if (g->n > 0 && !g->offsets) return false;
  if (g->m > 0 && (!g->indices || g->n == 0)) return false;
  return true;
}

// Layout of the jc_jaccard_ex workspace.
struct Layout {
  size_t engine, rows, s64, exact, work32, wi, ws, total;
};

Layout layout(int32_t n, int32_t m, bool weighted, bool ec) {
  Layout L{};
  size_t o = 0;
  L.engine = o;
  o += al(jc::engine_workspace_bytes(n, m));
  L.rows = o;
  if (weighted) o += al(jc::rowsum_scratch_bytes(n));
  L.s64 = o;
  if (weighted) o += al((size_t)std::max(n, 1) * sizeof(double));
  L.exact = o;
  if (weighted) o += al(sizeof(jc::ExactInfo));
  L.work32 = o;
  if (!weighted && !ec) o += al((size_t)std::max(n, 1) * sizeof(float));
  L.wi = o;
  if (!ec) o += al((size_t)std::max(m, 1) * (weighted ? sizeof(double) : sizeof(float)));
  L.ws = o;
  if (!ec) o += al((size_t)std::max(m, 1) * (weighted ? sizeof(double) : sizeof(float)));
  L.total = o;
  return L;
}

// RAII for a stream-ordered allocation.
struct AsyncBuf {
  void* p = nullptr;
  hipStream_t s = nullptr;
  ~AsyncBuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
};

int run_jaccard(const jc_csr* g, const float* vw, float* result, bool ec, unsigned flags,
                unsigned char* ws, const jc_tuning* tuning, hipStream_t s) {
  const bool weighted = vw != nullptr;
  const bool fixup = (flags & JC_FLAG_ASSUME_SYMMETRIC) == 0;
  const Layout L = layout(g->n, g->m, weighted, ec);
  jc::EngineArgs a{};
  a.n = g->n;
  a.m = g->m;
  a.off = g->offsets;
  a.ind = g->indices;
  a.vw = vw;
  a.small_max = tuning ? tuning->small_max : 0;
  a.small_blocks_per_cu = tuning ? tuning->small_blocks_per_cu : 0;
  a.wide_index = tuning ? (tuning->wide_index != 0) : 0;
  a.bitmap_blocks_per_cu = tuning ? tuning->bitmap_blocks_per_cu : 0;
  a.hash_table = tuning ? tuning->hash_table : 0;
  a.quot_max_disp = tuning ? tuning->quot_max_disp : 0;
  double* S64 = weighted ? reinterpret_cast<double*>(ws + L.s64) : nullptr;
  if (weighted) {
    jc::ExactInfo* ex = reinterpret_cast<jc::ExactInfo*>(ws + L.exact);
    JC_TRY(jc::launch_rowsum_f64(g->n, g->offsets, g->indices, vw, S64, ws + L.rows, s));
    JC_TRY(jc::launch_exact_info(g->n, vw, S64, ex, s));
    a.exact = ex;
  }
  a.S64 = S64;
  if (ec) {
    a.out32 = result;
    if (fixup) JC_TRY(jc::launch_fill_u32(reinterpret_cast<uint32_t*>(result), g->m, jc::kSentinel32, s));
    JC_TRY(jc::run_engine(weighted ? jc::kModeEcW : jc::kModeEcU, weighted, a, ws + L.engine, fixup, s));
    return JC_OK;
  }
  if (weighted) {
    double* wi = reinterpret_cast<double*>(ws + L.wi);
    double* wsum = reinterpret_cast<double*>(ws + L.ws);
    a.wi64 = wi;
    a.ws64 = wsum;
    if (fixup) JC_TRY(jc::launch_fill_u64(reinterpret_cast<uint64_t*>(wsum), g->m, jc::kSentinel64, s));
    JC_TRY(jc::run_engine(jc::kModeIsF64, true, a, ws + L.engine, fixup, s));
    JC_TRY(jc::launch_jw_f64(g->m, wi, wsum, result, s));
    return JC_OK;
  }
  float* work = reinterpret_cast<float*>(ws + L.work32);
  float* wi = reinterpret_cast<float*>(ws + L.wi);
  float* wsum = reinterpret_cast<float*>(ws + L.ws);
  JC_TRY(jc::launch_rowsum_f32(g->n, g->offsets, g->indices, nullptr, work, nullptr, s));
  a.work32 = work;
  a.out32 = wi;
  a.ws32 = wsum;
  a.accumulate = 0;
  if (fixup) JC_TRY(jc::launch_fill_u32(reinterpret_cast<uint32_t*>(wsum), g->m, jc::kSentinel32, s));
  JC_TRY(jc::run_engine(jc::kModeIsF32, false, a, ws + L.engine, fixup, s));
  JC_TRY(jc::launch_jw_f32(g->m, wi, wsum, result, s));
  return JC_OK;
}

// Scratch of jc_jaccard_noalloc: one array per device, held until the process
// ends. In host code `scratch` names a host-side stand-in, never the device
// memory: the device address comes from hipGetSymbolAddress, once per device
// (with that device current), and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[JC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (resolved on the
// first call on that device), or an error.
hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

namespace jc {
// Lets the other translation units report a failing HIP call through jc_last_hip_error().
int note_hip_error(hipError_t e) { return hip_fail(e); }
}  // namespace jc

extern "C" {

size_t jc_jaccard_workspace_size(int32_t n, int32_t m, int weighted, int edge_centric) {
  if (n < 0 || m < 0) return 0;
  return layout(n, m, weighted != 0, edge_centric != 0).total;
}

int jc_jaccard_tuned(const jc_csr* g, const float* vertex_weights, float* result, int edge_centric,
                     unsigned flags, void* workspace, size_t workspace_bytes, const jc_tuning* tuning,
                     jc_stream_t stream) {
  if (!csr_args_ok(g) || (g->m > 0 && !result) || (flags & ~JC_FLAG_ASSUME_SYMMETRIC)) return JC_ERR_INVALID_ARG;
  if (g->n == 0 || g->m == 0) return JC_OK;
  const bool ec = edge_centric != 0;
  const size_t need = layout(g->n, g->m, vertex_weights != nullptr, ec).total;
  hipStream_t s = hs(stream);
  if (workspace) {
    if (workspace_bytes < need) return JC_ERR_WORKSPACE;
    return run_jaccard(g, vertex_weights, result, ec, flags, static_cast<unsigned char*>(workspace), tuning, s);
  }
  AsyncBuf buf;
  buf.s = s;
  JC_TRY(hipMallocAsync(&buf.p, need, s));
  return run_jaccard(g, vertex_weights, result, ec, flags, static_cast<unsigned char*>(buf.p), tuning, s);
}

int jc_jaccard_ex(const jc_csr* g, const float* vertex_weights, float* result, int edge_centric,
                  unsigned flags, void* workspace, size_t workspace_bytes, jc_stream_t stream) {
  return jc_jaccard_tuned(g, vertex_weights, result, edge_centric, flags, workspace, workspace_bytes, nullptr,
                          stream);
}

int jc_jaccard(const jc_csr* g, const float* vertex_weights, float* result, int edge_centric,
               jc_stream_t stream) {
  return jc_jaccard_ex(g, vertex_weights, result, edge_centric, 0u, nullptr, 0, stream);
}

int jc_jaccard_noalloc(const jc_csr* g, const float* vertex_weights, float* result, int edge_centric,
                       unsigned flags, jc_stream_t stream) {
  if (!csr_args_ok(g) || (g->m > 0 && !result) || (flags & ~JC_FLAG_ASSUME_SYMMETRIC)) return JC_ERR_INVALID_ARG;
  if (g->n == 0 || g->m == 0) return JC_OK;
  const bool ec = edge_centric != 0;
  if (layout(g->n, g->m, vertex_weights != nullptr, ec).total > (size_t)JC_NOALLOC_SCRATCH_BYTES)
    return JC_ERR_SCRATCH_LIMIT;
  void* ws = nullptr;
  JC_TRY(scratch_address(&ws));
  // Every part of the workspace that must start at zero (schedule counters,
  // range facts, row-sum queue) is cleared by the kernels of the call itself.
  return run_jaccard(g, vertex_weights, result, ec, flags, static_cast<unsigned char*>(ws), nullptr, hs(stream));
}

int jc_jaccard_list(const jc_csr* g, const float* vertex_weights, int32_t num_pairs,
                    const int32_t* first, const int32_t* second, float* result, jc_stream_t stream) {
  if (!csr_args_ok(g) || num_pairs < 0) return JC_ERR_INVALID_ARG;
  if (num_pairs == 0) return JC_OK;
  if (!first || !second || !result || g->n == 0) return JC_ERR_INVALID_ARG;
  const bool weighted = vertex_weights != nullptr;
  JC_TRY(jc::run_pairs(weighted ? jc::kModeEcW : jc::kModeEcU, weighted, g->n, g->offsets, g->indices,
                       vertex_weights, nullptr, num_pairs, first, second, result, nullptr, 0, hs(stream)));
  return JC_OK;
}

int jc_vertex_weights(const jc_csr* g, float* W, jc_stream_t stream) {
  if (!csr_args_ok(g) || (g->n > 0 && !W) || (g->m > 0 && !g->edge_w)) return JC_ERR_INVALID_ARG;
  if (g->n == 0) return JC_OK;
  if (g->m == 0) {
    JC_TRY(jc::launch_fill_f32(W, g->n, 0.0f, hs(stream)));
    return JC_OK;
  }
  JC_TRY(jc::launch_vertex_weights(g->n, g->offsets, g->edge_w, W, nullptr, hs(stream)));
  return JC_OK;
}

size_t jc_vertex_weights_workspace_size(int32_t n) { return n < 0 ? 0 : al(jc::rowsum_scratch_bytes(n)); }

int jc_vertex_weights_ex(const jc_csr* g, float* W, void* workspace, size_t workspace_bytes, jc_stream_t stream) {
  if (!csr_args_ok(g) || (g->n > 0 && !W) || (g->m > 0 && !g->edge_w) || !workspace) return JC_ERR_INVALID_ARG;
  if (g->n == 0) return JC_OK;
  if (workspace_bytes < jc_vertex_weights_workspace_size(g->n)) return JC_ERR_WORKSPACE;
  if (g->m == 0) {
    JC_TRY(jc::launch_fill_f32(W, g->n, 0.0f, hs(stream)));
    return JC_OK;
  }
  JC_TRY(jc::launch_vertex_weights(g->n, g->offsets, g->edge_w, W, workspace, hs(stream)));
  return JC_OK;
}

int jc_nonzero_count(int32_t m, const float* result, float thresh, int64_t* count, jc_stream_t stream) {
  if (m < 0 || !count || (m > 0 && !result)) return JC_ERR_INVALID_ARG;
  *count = 0;
  if (m == 0) return JC_OK;
  hipStream_t s = hs(stream);
  AsyncBuf buf;
  buf.s = s;
  JC_TRY(hipMallocAsync(&buf.p, sizeof(unsigned long long), s));
  auto* c = static_cast<unsigned long long*>(buf.p);
  JC_TRY(jc::launch_fill_u64(reinterpret_cast<uint64_t*>(c), 1, 0ull, s));
  JC_TRY(jc::launch_count_above(m, result, thresh, c, s));
  unsigned long long h = 0;
  JC_TRY(hipMemcpyAsync(&h, c, sizeof(h), hipMemcpyDeviceToHost, s));
  JC_TRY(hipStreamSynchronize(s));
  *count = (int64_t)h;
  return JC_OK;
}

int jc_fill(int64_t n, float* x, float value, jc_stream_t stream) {
  if (n < 0 || (n > 0 && !x)) return JC_ERR_INVALID_ARG;
  JC_TRY(jc::launch_fill_f32(x, n, value, hs(stream)));
  return JC_OK;
}

int jc_rowsum(int32_t n, int32_t m, const int32_t* offsets, const int32_t* indices, const float* v,
              float* work, jc_stream_t stream) {
  if (n < 0 || m < 0 || (n > 0 && (!offsets || !work)) || (m > 0 && v && !indices)) return JC_ERR_INVALID_ARG;
  JC_TRY(jc::launch_rowsum_f32(n, offsets, indices, v, work, nullptr, hs(stream)));
  return JC_OK;
}

int jc_intersect(int32_t n, int32_t m, const int32_t* offsets, const int32_t* indices, const float* v,
                 const float* work, float* weight_i, float* weight_s, jc_stream_t stream) {
  jc_csr g{n, m, offsets, indices, nullptr};
  if (!csr_args_ok(&g) || (m > 0 && (!work || !weight_i || !weight_s))) return JC_ERR_INVALID_ARG;
  if (n == 0 || m == 0) return JC_OK;
  hipStream_t s = hs(stream);
  AsyncBuf buf;
  buf.s = s;
  JC_TRY(hipMallocAsync(&buf.p, jc::engine_workspace_bytes(n, m), s));
  jc::EngineArgs a{};
  a.n = n;
  a.m = m;
  a.off = offsets;
  a.ind = indices;
  a.vw = v;
  a.work32 = work;
  a.out32 = weight_i;
  a.ws32 = weight_s;
  a.accumulate = 1;
  a.small_max = 0;
  a.small_blocks_per_cu = 0;
  JC_TRY(jc::launch_fill_u32(reinterpret_cast<uint32_t*>(weight_s), m, jc::kSentinel32, s));
  JC_TRY(jc::run_engine(jc::kModeIsF32, v != nullptr, a, buf.p, true, s));
  return JC_OK;
}

int jc_intersect_pairs(int32_t n, int32_t num_pairs, const int32_t* offsets, const int32_t* indices,
                       const int32_t* first, const int32_t* second, const float* v, const float* work,
                       float* weight_i, float* weight_s, jc_stream_t stream) {
  if (n < 0 || num_pairs < 0 || (n > 0 && !offsets)) return JC_ERR_INVALID_ARG;
  if (num_pairs == 0) return JC_OK;
  if (n == 0 || !indices || !first || !second || !work || !weight_i || !weight_s) return JC_ERR_INVALID_ARG;
  JC_TRY(jc::run_pairs(jc::kModeIsF32, v != nullptr, n, offsets, indices, v, work, num_pairs, first, second,
                       weight_i, weight_s, 1, hs(stream)));
  return JC_OK;
}

int jc_jw(int32_t m, const float* weight_i, const float* weight_s, float* weight_j, jc_stream_t stream) {
  if (m < 0 || (m > 0 && (!weight_i || !weight_s || !weight_j))) return JC_ERR_INVALID_ARG;
  JC_TRY(jc::launch_jw_f32(m, weight_i, weight_s, weight_j, hs(stream)));
  return JC_OK;
}

int jc_ec_scan(int32_t m, int32_t n, const int32_t* offsets, int32_t* dest_ind, float* weight_j,
               jc_stream_t stream) {
  if (m < 0 || n < 0 || (n > 0 && !offsets) || (m > 0 && (!dest_ind || !weight_j))) return JC_ERR_INVALID_ARG;
  JC_TRY(jc::launch_ec_scan(n, offsets, dest_ind, weight_j, hs(stream)));
  return JC_OK;
}

int jc_ec_unweighted(int32_t m, int32_t n, const int32_t* offsets, const int32_t* indices,
                     const int32_t* dest_ind, float* weight_j, jc_stream_t stream) {
  (void)dest_ind;
  jc_csr g{n, m, offsets, indices, nullptr};
  return jc_jaccard_ex(&g, nullptr, weight_j, JC_EDGE_CENTRIC, 0u, nullptr, 0, stream);
}

int jc_weighted_sum_mode(const jc_csr* g, const float* vertex_weights, int* mode, jc_stream_t stream) {
  if (!csr_args_ok(g) || !vertex_weights || !mode) return JC_ERR_INVALID_ARG;
  *mode = JC_SUM_FP64;
  if (g->n == 0) return JC_OK;
  hipStream_t s = hs(stream);
  // the row sums and range facts exactly as run_jaccard computes them
  const size_t rows_b = al(jc::rowsum_scratch_bytes(g->n));
  const size_t s64_b = al((size_t)g->n * sizeof(double));
  AsyncBuf buf;
  buf.s = s;
  JC_TRY(hipMallocAsync(&buf.p, rows_b + s64_b + al(sizeof(jc::ExactInfo)), s));
  unsigned char* p = static_cast<unsigned char*>(buf.p);
  double* S64 = reinterpret_cast<double*>(p + rows_b);
  jc::ExactInfo* ex = reinterpret_cast<jc::ExactInfo*>(p + rows_b + s64_b);
  JC_TRY(jc::launch_rowsum_f64(g->n, g->offsets, g->indices, vertex_weights, S64, p, s));
  JC_TRY(jc::launch_exact_info(g->n, vertex_weights, S64, ex, s));
  jc::ExactInfo h{};
  JC_TRY(hipMemcpyAsync(&h, ex, sizeof(h), hipMemcpyDeviceToHost, s));
  JC_TRY(hipStreamSynchronize(s));
  *mode = jc::exact_sums_fit_f64(h) ? JC_SUM_FP64 : JC_SUM_DOUBLE_DOUBLE;
  return JC_OK;
}

int jc_degree_stats(const jc_csr* g, jc_graph_stats* stats, jc_stream_t stream) {
  if (!csr_args_ok(g) || !stats) return JC_ERR_INVALID_ARG;
  *stats = jc_graph_stats{};
  jc::bitmap_table_bytes(stats->bitmap_table_bytes);
  if (g->n == 0) return JC_OK;
  hipStream_t s = hs(stream);
  constexpr int kWords = 69;  // see launch_degree_stats
  AsyncBuf buf;
  buf.s = s;
  JC_TRY(hipMallocAsync(&buf.p, kWords * sizeof(unsigned long long), s));
  auto* d = static_cast<unsigned long long*>(buf.p);
  JC_TRY(jc::launch_fill_u64(reinterpret_cast<uint64_t*>(d), kWords, 0ull, s));
  JC_TRY(jc::launch_degree_stats(g->n, g->offsets, g->indices, d, s));
  unsigned long long h[kWords];
  JC_TRY(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s));
  JC_TRY(hipStreamSynchronize(s));
  for (int b = 0; b < 32; ++b) {
    stats->rows[b] = (int64_t)h[b];
    stats->entries[b] = (int64_t)h[32 + b];
  }
  stats->isolated = (int64_t)h[64];
  stats->long_rows = (int64_t)h[65];
  stats->long_bitmap_rows[0] = (int64_t)h[66];
  stats->long_bitmap_rows[1] = (int64_t)h[67];
  stats->max_degree = (int32_t)(uint32_t)h[68];
  return JC_OK;
}

int jc_config_from_stats(const jc_graph_stats* st, int32_t n, int weighted, int* edge_centric,
                         jc_tuning* tuning) {
  if (!st || !edge_centric || !tuning || n < 0) return JC_ERR_INVALID_ARG;
  (void)weighted;  // the measured choices are the same for both
  *edge_centric = JC_EDGE_CENTRIC;
  *tuning = jc_tuning{};
  tuning->small_max = jc::kSmallMax;
  tuning->small_blocks_per_cu = 32;
  tuning->bitmap_blocks_per_cu = 1;
  tuning->hash_table = n >= jc::kQuotMaxN ? 1 : 0;
  return JC_OK;
}

int jc_select_config(const jc_csr* g, int weighted, int* edge_centric, jc_tuning* tuning, jc_stream_t stream) {
  if (!csr_args_ok(g) || !edge_centric || !tuning) return JC_ERR_INVALID_ARG;
  jc_graph_stats st;
  const int rc = jc_degree_stats(g, &st, stream);
  if (rc != JC_OK) return rc;
  return jc_config_from_stats(&st, g->n, weighted, edge_centric, tuning);
}

int jc_validate_csr(const jc_csr* g, int* what, jc_stream_t stream) {
  if (what) *what = 0;
  if (!csr_args_ok(g)) return JC_ERR_INVALID_ARG;
  if (g->n == 0) return g->m == 0 ? JC_OK : JC_ERR_INVALID_CSR;
  hipStream_t s = hs(stream);
  AsyncBuf buf;
  buf.s = s;
  JC_TRY(hipMallocAsync(&buf.p, 4 * sizeof(int), s));
  int* f = static_cast<int*>(buf.p);
  JC_TRY(jc::launch_fill_u32(reinterpret_cast<uint32_t*>(f), 4, 0u, s));
  JC_TRY(jc::launch_validate(g->n, g->m, g->offsets, g->indices, f, s));
  int h[4] = {0, 0, 0, 0};
  JC_TRY(hipMemcpyAsync(h, f, sizeof(h), hipMemcpyDeviceToHost, s));
  JC_TRY(hipStreamSynchronize(s));
  if (h[0] == 0 && g->m > 0) {
    JC_TRY(jc::launch_validate_rows(g->n, g->offsets, g->indices, f, s));
    JC_TRY(hipMemcpyAsync(h, f, sizeof(h), hipMemcpyDeviceToHost, s));
    JC_TRY(hipStreamSynchronize(s));
  }
  const int mask = (h[0] ? 1 : 0) | (h[1] ? 2 : 0) | (h[2] ? 4 : 0);
  if (what) *what = mask;
  return mask ? JC_ERR_INVALID_CSR : JC_OK;
}

int jc_csr_is_symmetric(const jc_csr* g, int* symmetric, jc_stream_t stream) {
  if (!csr_args_ok(g) || !symmetric) return JC_ERR_INVALID_ARG;
  *symmetric = 1;
  if (g->n == 0 || g->m == 0) return JC_OK;
  hipStream_t s = hs(stream);
  AsyncBuf buf;
  buf.s = s;
  JC_TRY(hipMallocAsync(&buf.p, sizeof(int), s));
  int* f = static_cast<int*>(buf.p);
  JC_TRY(jc::launch_fill_u32(reinterpret_cast<uint32_t*>(f), 1, 0u, s));
  JC_TRY(jc::launch_symcheck(g->n, g->m, g->offsets, g->indices, f, s));
  int h = 0;
  JC_TRY(hipMemcpyAsync(&h, f, sizeof(h), hipMemcpyDeviceToHost, s));
  JC_TRY(hipStreamSynchronize(s));
  *symmetric = h ? 0 : 1;
  return JC_OK;
}

const char* jc_strerror(int code) {
  switch (code) {
    case JC_OK: return "success";
    case JC_ERR_INVALID_ARG: return "invalid argument";
    case JC_ERR_INVALID_CSR: return "invalid CSR graph";
    case JC_ERR_OUT_OF_MEMORY: return "device out of memory";
    case JC_ERR_HIP: return "HIP runtime error";
    case JC_ERR_WORKSPACE: return "workspace too small";
    case JC_ERR_SCRATCH_LIMIT: return "graph too large for the allocation-free entry point's fixed scratch";
    default: return "unknown error";
  }
}

int jc_last_hip_error(void) { return g_last_hip; }

const char* jc_version(void) { return "jaccard_amd 0.1.0 gfx950"; }

}  // extern "C"
