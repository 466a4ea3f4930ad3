// Owner engine: dispatch over the result modes, workspace size and the pair
// path. The kernels and the launch sequence are in jc_engine.hpp; each
// result mode is instantiated in a translation unit of its own
// (jc_engine_<mode>.hip), so the instantiations compile in parallel.
#include "jc_engine.hpp"

namespace jc {

extern template hipError_t run_engine_t<kModeEcU, false>(const EngineArgs&, void*, bool, hipStream_t);
extern template hipError_t run_engine_t<kModeEcW, true>(const EngineArgs&, void*, bool, hipStream_t);
extern template hipError_t run_engine_t<kModeIsF32, false>(const EngineArgs&, void*, bool, hipStream_t);
extern template hipError_t run_engine_t<kModeIsF32, true>(const EngineArgs&, void*, bool, hipStream_t);
extern template hipError_t run_engine_t<kModeIsF64, true>(const EngineArgs&, void*, bool, hipStream_t);

size_t engine_workspace_bytes(int32_t n, int64_t m) {
  size_t b = align_up(sizeof(Sched), 256);
  b += align_up((size_t)max_tasks(n, m) * sizeof(int2), 256);
  b += align_up((size_t)std::max<int32_t>(n, 1) * sizeof(int32_t), 256);
  if (JC_QTABLE && n < kQuotMaxN) b += align_up((size_t)max_tasks(n, m) * sizeof(int), 256);  // redo list
  return b;
}

hipError_t run_engine(int mode, bool weighted, const EngineArgs& base, void* ws, bool fixup,
                      hipStream_t stream) {
  switch (mode) {
    case kModeEcU: return run_engine_t<kModeEcU, false>(base, ws, fixup, stream);
    case kModeEcW: return run_engine_t<kModeEcW, true>(base, ws, fixup, stream);
    case kModeIsF32:
      return weighted ? run_engine_t<kModeIsF32, true>(base, ws, fixup, stream)
                      : run_engine_t<kModeIsF32, false>(base, ws, fixup, stream);
    case kModeIsF64: return run_engine_t<kModeIsF64, true>(base, ws, fixup, stream);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Degree statistics (jc_degree_stats): one thread per row, counts kept in LDS
// per workgroup and added to the global totals once. Layout of out: rows[32],
// entries[32], isolated, long rows, long rows whose bitmap fits (unweighted,
// weighted table layout), then the largest degree as the low word of out[68].
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_degree_stats(int32_t n, const int32_t* __restrict__ off,
                                                        const int32_t* __restrict__ ind, size_t bud_u,
                                                        size_t bud_w, unsigned long long* out) {
  __shared__ unsigned long long s[68];
  __shared__ int s_max;
  for (int i = threadIdx.x; i < 68; i += BLOCK) s[i] = 0ull;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  int dmax = 0;
  for (int32_t v = blockIdx.x * BLOCK + threadIdx.x; v < n; v += gridDim.x * BLOCK) {
    const int o = off[v], d = off[v + 1] - o;
    if (d <= 0) {
      atomicAdd(&s[64], 1ull);
      continue;
    }
    const int b = 31 - __clz(d);
    atomicAdd(&s[b], 1ull);
    atomicAdd(&s[32 + b], (unsigned long long)d);
    dmax = max(dmax, d);
    if (d > kSmallMax) {
      atomicAdd(&s[65], 1ull);
      const int32_t lo = ind[o], hi = ind[o + d - 1];
      if (bitmap_fits<false>(lo, hi, d, bud_u)) atomicAdd(&s[66], 1ull);
      if (bitmap_fits<true>(lo, hi, d, bud_w)) atomicAdd(&s[67], 1ull);
    }
  }
  atomicMax(&s_max, dmax);
  __syncthreads();
  for (int i = threadIdx.x; i < 68; i += BLOCK)
    if (s[i]) atomicAdd(&out[i], s[i]);
  if (threadIdx.x == 0) atomicMax(reinterpret_cast<int*>(&out[68]), s_max);
}

void bitmap_table_bytes(int64_t out[2]) {
  out[0] = (int64_t)BlockCfg<false>::bm_budget1;
  out[1] = (int64_t)BlockCfg<true>::bm_budget1;
}

hipError_t launch_degree_stats(int32_t n, const int32_t* off, const int32_t* ind, unsigned long long* out,
                               hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  constexpr int B = 256;
  // the bitmap budgets of the workgroup kernel with one workgroup per CU
  const size_t bud_u = BlockCfg<false>::bm_budget1, bud_w = BlockCfg<true>::bm_budget1;
  const int g = (int)std::min<int64_t>((n + B - 1) / B, (int64_t)device_cu_count() * 8);
  hipLaunchKernelGGL((k_degree_stats<B>), dim3(g), dim3(B), 0, stream, n, off, ind, bud_u, bud_w, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pair list: one wave per pair.
template <int MODE, bool W, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_pairs(int32_t n, const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ ind,
                                                 const float* __restrict__ vw,
                                                 const float* __restrict__ work32, int64_t np,
                                                 const int32_t* __restrict__ first,
                                                 const int32_t* __restrict__ second, float* out32,
                                                 float* ws32, int accumulate) {
  const int64_t nwaves = (int64_t)gridDim.x * (BLOCK / kWave);
  for (int64_t i = (int64_t)blockIdx.x * (BLOCK / kWave) + threadIdx.x / kWave; i < np; i += nwaves) {
    const int32_t pa = first[i], pb = second[i];
    if (pa < 0 || pa >= n || pb < 0 || pb >= n) {
      if (lane_id() == 0) {
        out32[i] = __int_as_float(0x7fc00000);
        if (MODE == kModeIsF32) ws32[i] = __int_as_float(0x7fc00000);
      }
      continue;
    }
    uint32_t cnt;
    DD acc;
    wave_pair<W>(off, ind, vw, pa, pb, cnt, acc);
    if constexpr (MODE == kModeEcU) {
      if (lane_id() == 0) {
        const int da = off[pa + 1] - off[pa], db = off[pb + 1] - off[pb];
        const float fI = (float)cnt;
        out32[i] = fI / ((float)(da + db) - fI);
      }
    } else if constexpr (MODE == kModeEcW) {
      const double sa = wave_row_wsum(off, ind, vw, pa);
      const double sb = wave_row_wsum(off, ind, vw, pb);
      if (lane_id() == 0) out32[i] = (float)(acc.hi / ((sa + sb) - acc.hi));
    } else {
      if (lane_id() == 0) {
        const float add = W ? (float)acc.hi : (float)cnt;
        out32[i] = (accumulate ? out32[i] : 0.0f) + add;
        ws32[i] = work32[pa] + work32[pb];
      }
    }
  }
}

hipError_t run_pairs(int mode, bool weighted, int32_t n, const int32_t* off, const int32_t* ind,
                     const float* vw, const float* work32, int64_t np, const int32_t* first,
                     const int32_t* second, float* out32, float* ws32, int accumulate,
                     hipStream_t stream) {
  if (np <= 0) return hipSuccess;
  constexpr int B = 256;
  const int g = (int)std::min<int64_t>((np + (B / kWave) - 1) / (B / kWave), (int64_t)device_cu_count() * 16);
  ProfScope ps("pairs", stream);
  switch (mode) {
    case kModeEcU:
      hipLaunchKernelGGL((k_pairs<kModeEcU, false, B>), dim3(g), dim3(B), 0, stream, n, off, ind, vw,
                         work32, np, first, second, out32, ws32, accumulate);
      break;
    case kModeEcW:
      hipLaunchKernelGGL((k_pairs<kModeEcW, true, B>), dim3(g), dim3(B), 0, stream, n, off, ind, vw,
                         work32, np, first, second, out32, ws32, accumulate);
      break;
    case kModeIsF32:
      if (weighted)
        hipLaunchKernelGGL((k_pairs<kModeIsF32, true, B>), dim3(g), dim3(B), 0, stream, n, off, ind,
                           vw, work32, np, first, second, out32, ws32, accumulate);
      else
        hipLaunchKernelGGL((k_pairs<kModeIsF32, false, B>), dim3(g), dim3(B), 0, stream, n, off, ind,
                           vw, work32, np, first, second, out32, ws32, accumulate);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace jc
