// Graph ingest on the device: the preparation the reference driver does on
// the host with std::set (main.cpp:98-120, readMtxToCSR.cpp:116-208) --
// optional reverse edges, self loops dropped, repeated (src, dst) merged,
// columns sorted -- as a radix sort of 64-bit (src, dst) keys (hipCUB), a
// flag/scan compaction and a row-boundary pass.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <algorithm>

#include "../../include/jaccard_amd.h"
#include "jc_internal.hpp"

namespace jc {
namespace {

// Keys (src << 32 | dst); self loops and ids outside [0, n) get key n << 32,
// which sorts after every real edge and is dropped. Reverse edges follow the
// originals, so a stable sort keeps an original's weight for a repeated pair.
__global__ void k_make_keys(int32_t n, int64_t e, const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                            const float* __restrict__ w, int sym, uint64_t* __restrict__ keys,
                            float* __restrict__ vals, int* bad) {
  const uint64_t drop = (uint64_t)(uint32_t)n << 32;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = src[i], v = dst[i];
    const bool ok = u >= 0 && u < n && v >= 0 && v < n;
    if (!ok) atomicOr(bad, 1);
    const bool keep = ok && u != v;
    keys[i] = keep ? ((uint64_t)(uint32_t)u << 32) | (uint32_t)v : drop;
    if (vals) vals[i] = w ? w[i] : 1.0f;
    if (sym) {
      keys[e + i] = keep ? ((uint64_t)(uint32_t)v << 32) | (uint32_t)u : drop;
      if (vals) vals[e + i] = w ? w[i] : 1.0f;
    }
  }
}

// flag[i] = sorted key i is a kept edge and differs from key i - 1
__global__ void k_mark(int64_t e2, uint64_t drop, const uint64_t* __restrict__ keys, int32_t* __restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e2; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    flag[i] = (k < drop && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
  }
}

__global__ void k_compact(int64_t e2, const uint64_t* __restrict__ keys, const float* __restrict__ vals,
                          const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int64_t capacity,
                          int32_t* __restrict__ csrc, int32_t* __restrict__ ind, float* __restrict__ wout) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e2; i += (int64_t)gridDim.x * blockDim.x) {
    if (!flag[i]) continue;
    const int64_t p = pos[i];
    if (p >= capacity) continue;  // reported to the caller after the count is read
    const uint64_t k = keys[i];
    csrc[p] = (int32_t)(k >> 32);
    ind[p] = (int32_t)(k & 0xFFFFFFFFu);
    if (wout) wout[p] = vals ? vals[i] : 1.0f;
  }
}

// offsets[r] = first kept edge of row r: entry p writes the rows between the
// previous entry's source and its own; p == m closes the rows up to n.
__global__ void k_row_bounds(int32_t n, int64_t m, const int32_t* __restrict__ csrc, int32_t* __restrict__ off) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p <= m; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = p < m ? csrc[p] : n;
    const int32_t up = p > 0 ? csrc[p - 1] : -1;
    for (int32_t r = up + 1; r <= u; ++r) off[r] = (int32_t)p;
  }
}

int grid_of(int64_t items) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, (int64_t)device_cu_count() * 32));
}

int ceil_log2_u64(uint64_t v) {
  int b = 0;
  while (b < 64 && (1ull << b) < v) ++b;
  return b;
}

}  // namespace
}  // namespace jc

extern "C" int jc_csr_from_edges(int32_t n, int64_t num_edges, const int32_t* src, const int32_t* dst,
                                 const float* edge_w, int symmetrize, int32_t* offsets, int32_t* indices,
                                 float* weights_out, int64_t capacity, int64_t* m_out, jc_stream_t stream) {
  if (n < 0 || num_edges < 0 || !offsets || !m_out || capacity < 0 || (num_edges > 0 && (!src || !dst)) ||
      (capacity > 0 && !indices))
    return JC_ERR_INVALID_ARG;
  const int64_t e2 = num_edges * (symmetrize ? 2 : 1);
  if (e2 >= INT32_MAX) return JC_ERR_INVALID_ARG;  // sort and scan index with int
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const bool weighted = weights_out != nullptr;
  // device temporaries: keys x2, (weights x2), flags, positions, source rows, counters
  size_t sort_bytes = 0, scan_bytes = 0;
  hipError_t err;
  if (weighted)
    err = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (float*)nullptr, (float*)nullptr, (int)e2, 0, 64, s);
  else
    err = hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr, (int)e2,
                                            0, 64, s);
  if (err != hipSuccess) return jc::note_hip_error(err);
  err = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, (int)e2, s);
  if (err != hipSuccess) return jc::note_hip_error(err);
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  const size_t n2 = (size_t)std::max<int64_t>(e2, 1);
  size_t o = 0;
  const size_t o_k0 = o; o += al(n2 * 8);
  const size_t o_k1 = o; o += al(n2 * 8);
  const size_t o_v0 = o; o += weighted ? al(n2 * 4) : 0;
  const size_t o_v1 = o; o += weighted ? al(n2 * 4) : 0;
  const size_t o_fl = o; o += al(n2 * 4);
  const size_t o_ps = o; o += al(n2 * 4);
  const size_t o_cs = o; o += al(n2 * 4);
  const size_t o_ct = o; o += al(4 * sizeof(int32_t));
  const size_t o_tmp = o; o += al(std::max(sort_bytes, scan_bytes));
  unsigned char* ws = nullptr;
  if ((err = hipMallocAsync(reinterpret_cast<void**>(&ws), o, s)) != hipSuccess) return jc::note_hip_error(err);
  uint64_t* k0 = reinterpret_cast<uint64_t*>(ws + o_k0);
  uint64_t* k1 = reinterpret_cast<uint64_t*>(ws + o_k1);
  float* v0 = weighted ? reinterpret_cast<float*>(ws + o_v0) : nullptr;
  float* v1 = weighted ? reinterpret_cast<float*>(ws + o_v1) : nullptr;
  int32_t* flag = reinterpret_cast<int32_t*>(ws + o_fl);
  int32_t* pos = reinterpret_cast<int32_t*>(ws + o_ps);
  int32_t* csrc = reinterpret_cast<int32_t*>(ws + o_cs);
  int32_t* ctr = reinterpret_cast<int32_t*>(ws + o_ct);  // [0] bad id, [1] last flag, [2] last position
  void* tmp = ws + o_tmp;
  int code = JC_OK;
  int32_t host[3] = {0, 0, 0};
  int64_t m = 0;
  do {
    if ((err = jc::launch_fill_u32(reinterpret_cast<uint32_t*>(ctr), 4, 0u, s)) != hipSuccess) break;
    if (num_edges > 0) {
      hipLaunchKernelGGL(jc::k_make_keys, dim3(jc::grid_of(num_edges)), dim3(256), 0, s, n, num_edges, src, dst,
                         edge_w, symmetrize, k0, v0, ctr);
      if ((err = hipGetLastError()) != hipSuccess) break;
      const int end_bit = 32 + std::max(1, jc::ceil_log2_u64((uint64_t)n + 1));
      size_t tb = sort_bytes;
      if (weighted)
        err = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, (int)e2, 0, std::min(64, end_bit), s);
      else
        err = hipcub::DeviceRadixSort::SortKeys(tmp, tb, k0, k1, (int)e2, 0, std::min(64, end_bit), s);
      if (err != hipSuccess) break;
      hipLaunchKernelGGL(jc::k_mark, dim3(jc::grid_of(e2)), dim3(256), 0, s, e2, (uint64_t)(uint32_t)n << 32, k1,
                         flag);
      if ((err = hipGetLastError()) != hipSuccess) break;
      tb = scan_bytes;
      if ((err = hipcub::DeviceScan::ExclusiveSum(tmp, tb, flag, pos, (int)e2, s)) != hipSuccess) break;
      if ((err = hipMemcpyAsync(&ctr[1], flag + (e2 - 1), sizeof(int32_t), hipMemcpyDeviceToDevice, s)) != hipSuccess)
        break;
      if ((err = hipMemcpyAsync(&ctr[2], pos + (e2 - 1), sizeof(int32_t), hipMemcpyDeviceToDevice, s)) != hipSuccess)
        break;
      hipLaunchKernelGGL(jc::k_compact, dim3(jc::grid_of(e2)), dim3(256), 0, s, e2, k1, v1, flag, pos, capacity,
                         csrc, indices, weights_out);
      if ((err = hipGetLastError()) != hipSuccess) break;
    }
    if ((err = hipMemcpyAsync(host, ctr, sizeof(host), hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
    m = num_edges > 0 ? (int64_t)host[1] + host[2] : 0;
    if (host[0]) {
      code = JC_ERR_INVALID_ARG;
      break;
    }
    if (m > capacity) {
      code = JC_ERR_WORKSPACE;
      break;
    }
    hipLaunchKernelGGL(jc::k_row_bounds, dim3(jc::grid_of(m + 1)), dim3(256), 0, s, n, m, csrc, offsets);
    if ((err = hipGetLastError()) != hipSuccess) break;
    if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
  } while (false);
  (void)hipFreeAsync(ws, s);
  if (err != hipSuccess) return jc::note_hip_error(err);
  *m_out = m;
  return code;
}
