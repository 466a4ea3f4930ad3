// Device-side building blocks shared by the Jaccard kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jc {

constexpr int kWave = 64;

// Weighted intersections split the long list into chunks of this many
// entries. The chunking fixes the fp64 summation order of every weighted
// coefficient (per-chunk lane sums, xor-butterfly, chunks added in ascending
// order), so it is one constant for every kernel that produces a weighted
// intersection: plain, batched, pair and fix-up paths give identical bits.
constexpr int kChunkW = 4096;
// Unweighted counts are exact integers, so their chunk size is free; a larger
// chunk means fewer chunk passes over hub rows.
constexpr int kChunkU = 8192;

// Lists up to this length are streamed whole (their clip range and mirror
// position come from the same pass); longer lists are clipped by wave
// searches first.
constexpr int kFullPassMax = 128;

// Sentinel written before an owner-engine pass; any position still holding it
// afterwards was not reached through a symmetric pair and is recomputed by
// the fix-up pass. Arithmetic here never produces this NaN payload.
constexpr uint32_t kSentinel32 = 0xFFFFFFFFu;
constexpr uint64_t kSentinel64 = 0xFFFFFFFFFFFFFFFFull;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// xor butterfly over the 64 lanes. fp addition commutes, so every lane ends
// with the same bits and the tree is fixed: the sum is deterministic.
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int s = 1; s < kWave; s <<= 1) v += __shfl_xor(v, s);
  return v;
}

__device__ __forceinline__ int wave_bcast(int v, int src) { return __shfl(v, src); }

// Wave-cooperative lower_bound on a sorted int32 array: the smallest i in
// [lo, hi) with a[i] >= key, or hi. All lanes must call it with the same
// arguments; each round narrows the range 64-fold (3 rounds cover 262144).
__device__ __forceinline__ int wave_lower_bound(const int32_t* __restrict__ a, int lo, int hi,
                                                int32_t key) {
  while (hi - lo > kWave) {
    const int stride = (hi - lo + kWave - 1) / kWave;
    const int idx = lo + lane_id() * stride;
    const bool lt = idx < hi && a[idx] < key;
    const int c = popc64(ballot(lt));
    if (c == 0) return lo;
    const int nlo = lo + (c - 1) * stride + 1;
    const int nhi = min(hi, lo + c * stride);
    lo = nlo;
    hi = nhi;
  }
  const int idx = lo + lane_id();
  const bool lt = idx < hi && a[idx] < key;
  return lo + popc64(ballot(lt));
}

// Per-lane lower_bound (independent keys per lane).
__device__ __forceinline__ int lane_lower_bound(const int32_t* __restrict__ a, int lo, int hi,
                                                int32_t key) {
  int len = hi - lo;
  while (len > 0) {
    const int half = len >> 1;
    if (a[lo + half] < key) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

// Multiplicative hash into a power-of-two table of 2^(32-shift) slots.
__device__ __forceinline__ uint32_t hslot(int32_t key, int shift) {
  return ((uint32_t)key * 0x9E3779B1u) >> shift;
}

__host__ __device__ __forceinline__ int ceil_log2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// Ownership of an undirected pair: the endpoint with the larger
// (degree, vertex id) computes it and writes both stored directions.
__device__ __forceinline__ bool owns(int32_t u, int du, int32_t x, int dx) {
  return dx < du || (dx == du && x < u);
}

}  // namespace jc
