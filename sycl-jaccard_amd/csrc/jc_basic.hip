// Streaming kernels of the vertex-centric pipeline and the EC row expansion.
//
//   k_fill             FillKernel            jaccard.cpp:150-182
//   k_rowsum_*         Jaccard_RowSumKernel  jaccard.cpp:235-313 (+ parallel_prefix_sum :53-137)
//   k_vertex_weights   weighted-VC preamble  main.cpp:211-275 (W(k) = sum of row k's edge weights)
//   k_jw_*             Jaccard_JwKernel      jaccard.cpp:607-653
//   k_ec_scan          Jaccard_ec_scan       jaccard.cpp:654-700
//   k_count_above      the "non zero pairs" count  jaccard.cpp:811-819
//
// Row sums are one wave per row: the row is read with coalesced 256-byte loads
// and reduced with an xor butterfly (double-double, exact for fp32 addends of
// moderate dynamic range) instead of the reference's LDS-emulated 32-lane
// shuffle scan with two barriers per step.
#include "jc_device.hpp"
#include "jc_internal.hpp"
#include "jc_profile.hpp"

namespace jc {

namespace {

struct DD2 {
  double hi, lo;
};

__device__ __forceinline__ DD2 add_dd(DD2 a, DD2 b) {
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (b.hi - bb);
  double t = a.lo + b.lo;
  t += e;
  const double hi = s + t;
  return DD2{hi, t - (hi - s)};
}

__device__ __forceinline__ DD2 add_d(DD2 a, double v) {
  const double s = a.hi + v;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (v - bb);
  const double t = a.lo + e;
  const double hi = s + t;
  return DD2{hi, t - (hi - s)};
}

int grid_for(int64_t items, int per_block, int cap_per_cu) {
  const int64_t g = (items + per_block - 1) / per_block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, (int64_t)device_cu_count() * cap_per_cu));
}

// Grid for kernels that stride over rows with one wave per row: an odd
// workgroup count, so rows at ids with many trailing zero bits (the hubs of
// RMAT-like graphs) do not all land on the same few waves.
int grid_rows(int64_t rows, int waves_per_block, int cap_per_cu) {
  return grid_for(rows, waves_per_block, cap_per_cu) | 1;
}

}  // namespace

template <typename T>
__global__ void k_fill(T* __restrict__ x, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

hipError_t launch_fill_f32(float* x, int64_t n, float v, hipStream_t s) {
  ProfScope ps_("fill", s);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n, 256 * 4, 32)), dim3(256), 0, s, x, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* x, int64_t n, uint32_t v, hipStream_t s) {
  ProfScope ps_("fill", s);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid_for(n, 256 * 4, 32)), dim3(256), 0, s, x, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_u64(uint64_t* x, int64_t n, uint64_t v, hipStream_t s) {
  ProfScope ps_("fill", s);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid_for(n, 256 * 4, 32)), dim3(256), 0, s, x, n, v);
  return hipGetLastError();
}

// Unweighted row sum: work[row] = (float)degree (jaccard.cpp:257-258).
__global__ void k_rowsum_deg(int32_t n, const int32_t* __restrict__ off, float* __restrict__ work) {
  for (int32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    work[u] = (float)(off[u + 1] - off[u]);
}

// Weighted row sum over column weights: sum_{k in row u} v[ind[k]] (ind may be
// nullptr: identity index, i.e. the row's own edge weights, as main.cpp:216-257
// arranges with its presumInd buffer). By row length: rows of at most
// kRowShort entries by groups of 4 lanes (k_rowsum_w), rows up to kRowWave
// entries one wave per row from a list (k_rowsum_w_med), longer rows one
// workgroup per row from a list (k_rowsum_w_long). Loads are unrolled 4x per
// lane.
#ifndef JC_ROW_WAVE
#define JC_ROW_WAVE 2048
#endif
constexpr int kRowWave = JC_ROW_WAVE;
constexpr int kRowU = 4;

// Loads of one unrolled step are unconditional (clamped index) so they are
// all in flight together; IDX selects v[ind[p]] (column weights) or v[p]
// (the row's own edge weights).
template <bool IDX, int U>
__device__ __forceinline__ void gather_step(const int32_t* __restrict__ ind, const float* __restrict__ v, int p0,
                                            int stride, int lane, int end, float (&x)[U]) {
  int q[U];
#pragma unroll
  for (int t = 0; t < U; ++t) q[t] = min(p0 + t * stride + lane, end - 1);
  if constexpr (IDX) {
#pragma unroll
    for (int t = 0; t < U; ++t) q[t] = ind[q[t]];
  }
#pragma unroll
  for (int t = 0; t < U; ++t) x[t] = v[q[t]];
#pragma unroll
  for (int t = 0; t < U; ++t)
    if (p0 + t * stride + lane >= end) x[t] = 0.0f;
}

// Rows of at most kRowShort entries (most rows of power-law graphs), 16
// consecutive rows per wave and iteration, each summed by its own group of 4
// lanes (groups of 8 measured the same or slightly slower). (One wave per row
// leaves most lanes idle and pays a row's load latency per row.) Rows of up
// to kRowWave entries are listed by k_collect_rows and summed one wave per
// row by k_rowsum_w_med: summed here one after another by the whole wave,
// they made the waves that met many of them (power-law graphs put them at
// structured ids) the kernel's tail (RMAT-22: gathered sums 1.52 -> 0.83 ms,
// vertex weights 0.63 -> 0.41 ms with the lists). The double-double sums are
// exact, so the lane split does not change the result.
constexpr int kRowShort = 64;
#ifndef JC_ROW_GROUP
#define JC_ROW_GROUP 4
#endif
constexpr int kRowGroup = JC_ROW_GROUP;

template <typename OutT, bool IDX>
__global__ void k_rowsum_w(int32_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ ind,
                           const float* __restrict__ v, OutT* __restrict__ out) {
  constexpr int R = kWave / kRowGroup;  // rows per wave and iteration
  const int lane = lane_id();
  const int gl = lane % kRowGroup;
  const int nw = gridDim.x * (blockDim.x / kWave);
  for (int32_t base = (blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * R; base < n; base += nw * R) {
    const int32_t u = base + lane / kRowGroup;
    int o = 0, d = 0;
    if (u < n) {
      o = off[u];
      d = off[u + 1] - o;
    }
    const bool short_row = u < n && d <= kRowShort;
    {
      const int ds = short_row ? d : 0;
      DD2 acc{0.0, 0.0};
      for (int p0 = o; p0 < o + ds; p0 += kRowGroup * kRowU) {
        float x[kRowU];
        gather_step<IDX, kRowU>(ind, v, p0, kRowGroup, gl, o + ds, x);
#pragma unroll
        for (int t = 0; t < kRowU; ++t) acc = add_d(acc, (double)x[t]);
      }
#pragma unroll
      for (int sh = 1; sh < kRowGroup; sh <<= 1)
        acc = add_dd(acc, DD2{__shfl_xor(acc.hi, sh), __shfl_xor(acc.lo, sh)});
      if (short_row && gl == 0) out[u] = (OutT)acc.hi;
    }
  }
}

// Rows of kRowShort + 1 .. kRowWave entries, one wave per row; wave w of the
// grid takes the list's entries w, w + waves, ..., and the grid has several
// times the waves the device holds at once, so the dispatcher balances the
// rows' unequal lengths (an atomic cursor shared by every wave serialises on
// its one address: measured slower). The list is stored from the end of
// `list` backwards (k_collect_rows).
#ifndef JC_MED_U
#define JC_MED_U 4
#endif
constexpr int kMedU = JC_MED_U;

template <typename OutT, bool IDX>
__global__ void k_rowsum_w_med(int32_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ ind,
                               const float* __restrict__ v, OutT* __restrict__ out, const int* __restrict__ list,
                               const int* __restrict__ counters) {
  const int lane = lane_id();
  const int count = counters[2];
  const int nw = gridDim.x * (blockDim.x / kWave);
  for (int r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < count; r += nw) {
    const int32_t u = list[n - 1 - r];
    const int o = off[u], d = off[u + 1] - o;
    DD2 acc{0.0, 0.0};
    for (int p0 = o; p0 < o + d; p0 += kWave * kMedU) {
      float x[kMedU];
      gather_step<IDX, kMedU>(ind, v, p0, kWave, lane, o + d, x);
#pragma unroll
      for (int t = 0; t < kMedU; ++t) acc = add_d(acc, (double)x[t]);
    }
#pragma unroll
    for (int sh = 1; sh < kWave; sh <<= 1) acc = add_dd(acc, DD2{__shfl_xor(acc.hi, sh), __shfl_xor(acc.lo, sh)});
    if (lane == 0) out[u] = (OutT)acc.hi;
  }
}

// Rows longer than kRowWave fill list[0 ..) (counters[0]), taken from an
// atomic cursor (counters[1]) by a persistent grid of 1024-thread
// workgroups, one row each. (A static row-to-workgroup map is not used: in
// power-law graphs the long rows sit at structured ids -- RMAT hubs have many
// trailing zero bits -- and any fixed stride piles them onto a few
// workgroups.) Rows of kRowShort + 1 .. kRowWave entries fill
// list[n-1] downwards (counters[2]) for k_rowsum_w_med. The lists are
// disjoint, so n entries hold both. Each workgroup takes a contiguous range
// of rows: it counts them, scans the counts over its threads and reserves
// its places with one atomic per list, then writes (one atomic per wave on
// one counter serialised: 0.22 ms on RMAT-22).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_collect_rows(int32_t n, const int32_t* __restrict__ off,
                                                        int* __restrict__ list, int* __restrict__ counters) {
  constexpr int NW = BLOCK / kWave;
  __shared__ int s_w[2][NW];
  __shared__ int s_base[2];
  const int lane = lane_id(), w = threadIdx.x / kWave;
  const int64_t per = ((int64_t)n + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = per * blockIdx.x, b1 = b0 + per;
  const int32_t r0 = (int32_t)(b0 < n ? b0 : n);
  const int32_t r1 = (int32_t)(b1 < n ? b1 : n);
  int cl = 0, cm = 0;
  for (int32_t u = r0 + threadIdx.x; u < r1; u += BLOCK) {
    const int d = off[u + 1] - off[u];
    cl += d > kRowWave ? 1 : 0;
    cm += d > kRowShort && d <= kRowWave ? 1 : 0;
  }
  // inclusive scans over the wave, then over the waves
  int il = cl, im = cm;
#pragma unroll
  for (int sh = 1; sh < kWave; sh <<= 1) {
    const int tl = __shfl_up(il, sh), tm = __shfl_up(im, sh);
    if (lane >= sh) {
      il += tl;
      im += tm;
    }
  }
  if (lane == kWave - 1) {
    s_w[0][w] = il;
    s_w[1][w] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int tl = 0, tm = 0;
    for (int k = 0; k < NW; ++k) {
      const int a = s_w[0][k], b = s_w[1][k];
      s_w[0][k] = tl;
      s_w[1][k] = tm;
      tl += a;
      tm += b;
    }
    s_base[0] = tl ? atomicAdd(&counters[0], tl) : 0;
    s_base[1] = tm ? atomicAdd(&counters[2], tm) : 0;
  }
  __syncthreads();
  int pl = s_base[0] + s_w[0][w] + il - cl;
  int pm = s_base[1] + s_w[1][w] + im - cm;
  for (int32_t u = r0 + threadIdx.x; u < r1; u += BLOCK) {
    const int d = off[u + 1] - off[u];
    if (d > kRowWave) list[pl++] = u;
    else if (d > kRowShort) list[n - 1 - pm++] = u;
  }
}

template <typename OutT, bool IDX, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_rowsum_w_long(const int32_t* __restrict__ off,
                                                         const int32_t* __restrict__ ind,
                                                         const float* __restrict__ v, OutT* __restrict__ out,
                                                         const int* __restrict__ list, int* counters) {
  __shared__ double part[2][BLOCK / kWave];
  __shared__ int s_row;
  const int count = counters[0];
  for (;;) {
    if (threadIdx.x == 0) {
      const int r = atomicAdd(&counters[1], 1);
      s_row = r < count ? list[r] : -1;
    }
    __syncthreads();
    const int32_t u = s_row;
    if (u < 0) return;
    const int o = off[u], d = off[u + 1] - o;
    DD2 acc{0.0, 0.0};
    for (int p0 = o; p0 < o + d; p0 += BLOCK * kRowU) {
      float x[kRowU];
      gather_step<IDX, kRowU>(ind, v, p0, BLOCK, threadIdx.x, o + d, x);
#pragma unroll
      for (int t = 0; t < kRowU; ++t) acc = add_d(acc, (double)x[t]);
    }
    for (int s = 1; s < kWave; s <<= 1) acc = add_dd(acc, DD2{__shfl_xor(acc.hi, s), __shfl_xor(acc.lo, s)});
    if (lane_id() == 0) {
      part[0][threadIdx.x / kWave] = acc.hi;
      part[1][threadIdx.x / kWave] = acc.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      DD2 t{0.0, 0.0};
      for (int w = 0; w < BLOCK / kWave; ++w) t = add_dd(t, DD2{part[0][w], part[1][w]});
      out[u] = (OutT)t.hi;
    }
    __syncthreads();
  }
}

size_t rowsum_scratch_bytes(int32_t n) { return ((size_t)std::max(n, 1) + 4) * sizeof(int); }

// scratch: rowsum_scratch_bytes(n) bytes, or nullptr to use a stream-ordered
// allocation for the call.
template <typename OutT>
static hipError_t rowsum_both(int32_t n, const int32_t* off, const int32_t* ind, const float* v, OutT* out,
                              void* scratch, hipStream_t s) {
  void* own = nullptr;
  hipError_t err;
  if (!scratch) {
    if ((err = hipMallocAsync(&own, rowsum_scratch_bytes(n), s)) != hipSuccess) return err;
    scratch = own;
  }
  int* counters = static_cast<int*>(scratch);
  int* list = counters + 4;
  const int cus = device_cu_count();
  hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(counters), (int64_t)4, 0u);
  hipLaunchKernelGGL(k_collect_rows<256>, dim3(grid_for(n, 256 * 4, 4)), dim3(256), 0, s, n, off, list, counters);
  const int gs = grid_rows((n + kWave / kRowGroup - 1) / (kWave / kRowGroup), 4, 64);
  const int gm = grid_rows(n, 4, 64);  // at most n listed rows, one per wave
  if (ind) {
    hipLaunchKernelGGL((k_rowsum_w<OutT, true>), dim3(gs), dim3(256), 0, s, n, off, ind, v, out);
    hipLaunchKernelGGL((k_rowsum_w_med<OutT, true>), dim3(gm), dim3(256), 0, s, n, off, ind, v, out, list,
                       counters);
    hipLaunchKernelGGL((k_rowsum_w_long<OutT, true, 1024>), dim3(cus * 2), dim3(1024), 0, s, off, ind, v, out,
                       list, counters);
  } else {
    hipLaunchKernelGGL((k_rowsum_w<OutT, false>), dim3(gs), dim3(256), 0, s, n, off, ind, v, out);
    hipLaunchKernelGGL((k_rowsum_w_med<OutT, false>), dim3(gm), dim3(256), 0, s, n, off, ind, v, out, list,
                       counters);
    hipLaunchKernelGGL((k_rowsum_w_long<OutT, false, 1024>), dim3(cus * 2), dim3(1024), 0, s, off, ind, v, out,
                       list, counters);
  }
  err = hipGetLastError();
  if (own) {
    hipError_t e2 = hipFreeAsync(own, s);
    if (err == hipSuccess) err = e2;
  }
  return err;
}

hipError_t launch_rowsum_f32(int32_t n, const int32_t* off, const int32_t* ind, const float* v,
                             float* work, void* scratch, hipStream_t s) {
  ProfScope ps_("rowsum", s);
  if (n <= 0) return hipSuccess;
  if (v == nullptr) {
    hipLaunchKernelGGL(k_rowsum_deg, dim3(grid_for(n, 256, 32)), dim3(256), 0, s, n, off, work);
    return hipGetLastError();
  }
  return rowsum_both<float>(n, off, ind, v, work, scratch, s);
}

hipError_t launch_rowsum_f64(int32_t n, const int32_t* off, const int32_t* ind, const float* v,
                             double* S, void* scratch, hipStream_t s) {
  ProfScope ps_("rowsum", s);
  if (n <= 0) return hipSuccess;
  return rowsum_both<double>(n, off, ind, v, S, scratch, s);
}

hipError_t launch_vertex_weights(int32_t n, const int32_t* off, const float* edge_w, float* W,
                                 void* scratch, hipStream_t s) {
  ProfScope ps_("vertex_weights", s);
  if (n <= 0) return hipSuccess;
  return rowsum_both<float>(n, off, nullptr, edge_w, W, scratch, s);
}

// Jaccard weight: weight_j = Wi / (Ws - Wi) (jaccard.cpp:614-619).
__global__ void k_jw_f32(int64_t m, const float* __restrict__ wi, const float* __restrict__ ws,
                         float* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const float Wi = wi[j];
    const float Ws = ws[j];
    out[j] = Wi / (Ws - Wi);
  }
}

__global__ void k_jw_f64(int64_t m, const double* __restrict__ wi, const double* __restrict__ ws,
                         float* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const double Wi = wi[j];
    const double Ws = ws[j];
    out[j] = (float)(Wi / (Ws - Wi));
  }
}

hipError_t launch_jw_f32(int64_t m, const float* wi, const float* ws, float* out, hipStream_t s) {
  ProfScope ps_("jw", s);
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jw_f32, dim3(grid_for(m, 256 * 4, 32)), dim3(256), 0, s, m, wi, ws, out);
  return hipGetLastError();
}

hipError_t launch_jw_f64(int64_t m, const double* wi, const double* ws, float* out, hipStream_t s) {
  ProfScope ps_("jw", s);
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jw_f64, dim3(grid_for(m, 256 * 4, 32)), dim3(256), 0, s, m, wi, ws, out);
  return hipGetLastError();
}

// Row expansion: dest[k] = row(k), weight_j[k] = 0. One wave per row, so the
// writes of a row are coalesced (the reference writes a row with one thread).
__global__ void k_ec_scan(int32_t n, const int32_t* __restrict__ off, int32_t* __restrict__ dest,
                          float* __restrict__ wj) {
  const int nw = gridDim.x * (blockDim.x / kWave);
  for (int32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += nw) {
    const int o = off[u], e = off[u + 1];
    for (int p = o + lane_id(); p < e; p += kWave) {
      dest[p] = u;
      wj[p] = 0.0f;
    }
  }
}

hipError_t launch_ec_scan(int32_t n, const int32_t* off, int32_t* dest, float* wj, hipStream_t s) {
  ProfScope ps_("ec_scan", s);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ec_scan, dim3(grid_rows(n, 4, 64)), dim3(256), 0, s, n, off, dest, wj);
  return hipGetLastError();
}

__global__ void k_count_above(int64_t m, const float* __restrict__ x, float thresh,
                              unsigned long long* cnt) {
  unsigned long long c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    c += x[i] > thresh ? 1u : 0u;
  for (int s = 1; s < kWave; s <<= 1) c += __shfl_xor(c, s);
  if (lane_id() == 0 && c) atomicAdd(cnt, c);
}

hipError_t launch_count_above(int64_t m, const float* x, float thresh, unsigned long long* cnt,
                              hipStream_t s) {
  ProfScope ps_("count_above", s);
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_above, dim3(grid_for(m, 256 * 8, 16)), dim3(256), 0, s, m, x, thresh, cnt);
  return hipGetLastError();
}

// CSR validation. flags[0]: offsets bad (offsets[0] != 0, decreasing, or
// offsets[n] != m, or out of [0, m]); flags[1]: a column out of [0, n);
// flags[2]: a row not strictly increasing (unsorted or duplicate columns).
// Every read is bounds-checked before it is made.
__global__ void k_validate(int32_t n, int64_t m, const int32_t* __restrict__ off,
                           const int32_t* __restrict__ ind, int* flags) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nt = (int64_t)gridDim.x * blockDim.x;
  if (tid == 0 && (off[0] != 0 || (int64_t)off[n] != m)) atomicOr(&flags[0], 1);
  for (int64_t u = tid; u < n; u += nt) {
    const int32_t a = off[u], b = off[u + 1];
    if (a < 0 || b < a || (int64_t)b > m) atomicOr(&flags[0], 1);
  }
  for (int64_t k = tid; k < m; k += nt) {
    const int32_t c = ind[k];
    if (c < 0 || c >= n) atomicOr(&flags[1], 1);
  }
}

// Second pass (only run when the offsets are valid): sortedness per row.
__global__ void k_validate_rows(int32_t n, const int32_t* __restrict__ off,
                                const int32_t* __restrict__ ind, int* flags) {
  const int nw = gridDim.x * (blockDim.x / kWave);
  for (int32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += nw) {
    const int o = off[u], e = off[u + 1];
    bool bad = false;
    for (int p = o + 1 + lane_id(); p < e; p += kWave) bad |= ind[p] <= ind[p - 1];
    if (ballot(bad) && lane_id() == 0) atomicOr(&flags[2], 1);
  }
}

hipError_t launch_validate(int32_t n, int64_t m, const int32_t* off, const int32_t* ind, int* flags,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_validate, dim3(grid_for(std::max<int64_t>(n, m), 256 * 4, 16)), dim3(256), 0, s, n, m,
                     off, ind, flags);
  return hipGetLastError();
}

hipError_t launch_validate_rows(int32_t n, const int32_t* off, const int32_t* ind, int* flags,
                                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_validate_rows, dim3(grid_rows(n, 4, 64)), dim3(256), 0, s, n, off, ind, flags);
  return hipGetLastError();
}

// Symmetry check: flag if some stored (u, v) has no stored (v, u).
__global__ void k_symcheck(int32_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ ind,
                           int* flag) {
  const int nw = gridDim.x * (blockDim.x / kWave);
  for (int32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += nw) {
    bool bad = false;
    for (int p = off[u] + lane_id(); p < off[u + 1]; p += kWave) {
      const int32_t v = ind[p];
      const int lo = off[v], hi = off[v + 1];
      const int q = lane_lower_bound(ind, lo, hi, u);
      bad |= !(q < hi && ind[q] == u);
    }
    if (ballot(bad) && lane_id() == 0) atomicOr(flag, 1);
  }
}

hipError_t launch_symcheck(int32_t n, int64_t m, const int32_t* off, const int32_t* ind, int* flag,
                           hipStream_t s) {
  (void)m;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_symcheck, dim3(grid_rows(n, 4, 64)), dim3(256), 0, s, n, off, ind, flag);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Range facts of vertex weights and row sums (ExactInfo): smallest exponent
// of a set bit over all weights, bound exponent of the largest row sum, and
// whether any weight or sum is negative or not finite.
__global__ void k_exact_init(ExactInfo* e) {
  if (threadIdx.x == 0) *e = ExactInfo{INT32_MAX, INT32_MIN, 0, 0};
}

__global__ __launch_bounds__(256) void k_exact_scan(int32_t n, const float* __restrict__ vw,
                                                    const double* __restrict__ S, ExactInfo* info) {
  int lsb = INT32_MAX, mx = INT32_MIN, bad = 0;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t b = __float_as_uint(vw[i]);
    const uint32_t ex = (b >> 23) & 0xFFu, man = b & 0x7FFFFFu;
    if ((b >> 31) && (b << 1)) bad = 1;  // negative (-0 is harmless)
    if (ex == 0xFFu) {
      bad = 1;
    } else if (ex != 0 || man != 0) {
      const uint32_t full = ex ? (man | 0x800000u) : man;
      lsb = min(lsb, (ex ? (int)ex - 150 : -149) + __builtin_ctz(full));
    }
    const double sv = S[i];
    if (!(sv >= 0.0) || isinf(sv))
      bad = 1;
    else if (sv > 0.0)
      mx = max(mx, ilogb(sv) + 1);
  }
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    lsb = min(lsb, __shfl_xor(lsb, o));
    mx = max(mx, __shfl_xor(mx, o));
    bad |= __shfl_xor(bad, o);
  }
  // one set of atomics per workgroup (per wave, the few counters serialise)
  __shared__ int s_r[3][256 / kWave];
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_r[0][w] = lsb;
    s_r[1][w] = mx;
    s_r[2][w] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x / kWave); ++k) {
      lsb = min(lsb, s_r[0][k]);
      mx = max(mx, s_r[1][k]);
      bad |= s_r[2][k];
    }
    if (lsb != INT32_MAX) atomicMin(&info->min_lsb, lsb);
    if (mx != INT32_MIN) atomicMax(&info->max_exp, mx);
    if (bad) atomicOr(&info->bad, 1);
  }
}

hipError_t launch_exact_info(int32_t n, const float* vw, const double* S, ExactInfo* info, hipStream_t s) {
  hipLaunchKernelGGL(k_exact_init, dim3(1), dim3(64), 0, s, info);
  if (n > 0) hipLaunchKernelGGL(k_exact_scan, dim3(grid_for(n, 256 * 4, 4)), dim3(256), 0, s, n, vw, S, info);
  return hipGetLastError();
}

int device_cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cached[dev] = v;
  }
  return cached[dev];
}

}  // namespace jc
