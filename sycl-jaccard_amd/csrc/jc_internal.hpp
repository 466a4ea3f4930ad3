// Internal interfaces between the C ABI layer and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace jc {

// Epilogue of an intersection: what is written for a pair (u, x).
enum Mode : int {
  kModeEcU = 0,   // EC unweighted: out = I / ((float)(du+dx) - I)
  kModeEcW = 1,   // EC weighted:   out = (float)(I / ((S[u]+S[x]) - I)), fp64
  kModeIsF32 = 2, // VC intersection, reference buffers: weight_i (+)= I, weight_s = work[u]+work[x]
  kModeIsF64 = 3, // VC intersection, fp64 internals: wi64 = I, ws64 = S[u]+S[x]
};

// Rows up to this degree are handled one wave per row; longer rows by a
// workgroup per task. Tasks cover at most kTaskSpan consecutive entries of a
// row (sized so item list, per-item state and a 128 KiB hash table fit the
// 160 KiB LDS of a CU, one 1024-thread workgroup per CU).
#ifndef JC_TASK_SPAN
#define JC_TASK_SPAN 896
#endif
constexpr int kSmallMax = 128;
constexpr int kTaskSpan = JC_TASK_SPAN;
// Graphs with fewer vertices than this use quotient hash tables (16-bit keys):
// every id is at most 2^24 - 2, so the padding key of a row's last pass
// (INT32_MAX, whose low 24 bits are 2^24 - 1) hashes like no id.
constexpr int32_t kQuotMaxN = 1 << 24;

// Range facts of the vertex weights of a weighted pass (launch_exact_info):
// every weight is a multiple of 2^min_lsb, every row sum is below 2^max_exp,
// bad != 0 if a weight is negative or not finite.
struct ExactInfo {
  int min_lsb;
  int max_exp;
  int bad;
  int pad;
};

// True when every weight sum of a pass is exact in plain fp64: all vertex
// weights non-negative multiples of 2^min_lsb and every row sum below
// 2^max_exp with max_exp - min_lsb <= 52, so every partial sum of a subset of
// a row fits 53 bits; double-double and plain fp64 then give the same bits.
// No non-zero weight (min_lsb unset) or no non-zero row sum (max_exp unset,
// so every weight that occurs in a row is 0) makes every sum 0: exact too.
// The difference is taken in 64 bits (the unset values are INT32_MAX and
// INT32_MIN).
__host__ __device__ inline bool exact_sums_fit_f64(const ExactInfo& e) {
  if (e.bad != 0) return false;
  if (e.min_lsb == INT32_MAX || e.max_exp == INT32_MIN) return true;
  return (int64_t)e.max_exp - (int64_t)e.min_lsb <= 52;
}

struct EngineArgs {
  int32_t n;
  int64_t m;
  const int32_t* off;
  const int32_t* ind;
  const float* vw;       // vertex weights (weighted modes), else nullptr
  const double* S64;     // fp64 row sums of vw (kModeEcW, kModeIsF64)
  const float* work32;   // fp32 row sums (kModeIsF32)
  float* out32;          // EC result, or weight_i (kModeIsF32)
  float* ws32;           // weight_s (kModeIsF32)
  double* wi64;          // kModeIsF64
  double* ws64;          // kModeIsF64
  int accumulate;        // kModeIsF32: add I to the prior weight_i contents
  // schedule (device memory)
  const int2* tasks;     // (row, piece) for rows with degree > kSmallMax, degree-descending buckets
  const int* num_tasks;
  int* task_counter;
  const int32_t* small_rows;  // rows with 1 <= degree <= small_max
  const int* num_small;
  // tuning (results do not depend on these)
  int small_max;              // degree bin threshold: one wave per row up to here (1..kSmallMax)
  int small_blocks_per_cu;    // grid of the one-wave-per-row kernel
  int wide_index;             // 1: rows are read with 64-bit addresses (set by the
                              // launcher when m * 4 does not fit 32-bit offsets)
  int bitmap_blocks_per_cu;   // workgroups per CU of the bitmap-task kernel (1 or 2)
  int hash_table;             // 1: 32-bit tables even where quotient tables apply
  int quot_max_disp;          // > 0: quotient keys move at most quot_max_disp - 1 buckets
  uint32_t bm_budget;         // bytes of the bitmap table region (set by the launcher)
  // Row segments (device memory, optional): when nseg > 0 only the rows of
  // the segments are computed. segs[i] = (first row, running row count before
  // segment i) for i < nseg, and segs[nseg] = (0, total rows); seg_rows and
  // seg_edges are the totals (rows and stored edges of all segments).
  const int2* segs;
  int nseg;
  int32_t seg_rows;
  int64_t seg_edges;
  // Weighted modes: optional range facts of vw and S64 (device memory); when
  // they show all sums are exact in fp64 the kernels add in plain fp64.
  const ExactInfo* exact;
  // Hash-table tasks the quotient-table kernel had to abandon (device memory,
  // set by the launcher), and whether a launch is the 32-bit redo of them.
  int* redo;
  int* num_redo;
  int redo_pass;
};

// Workspace needed by an owner-engine pass over a graph with n vertices and m
// stored edges (task list, small-row list, counters). With row segments, n and
// m are the segments' totals.
size_t engine_workspace_bytes(int32_t n, int64_t m);

// Owner engine: every pair written once per stored direction. When
// check_sentinel is set the target array was pre-filled with the sentinel and
// a fix-up pass recomputes positions that no symmetric pair reached.
// ws must hold engine_workspace_bytes(n, m) bytes.
hipError_t run_engine(int mode, bool weighted, const EngineArgs& base, void* ws, bool fixup,
                      hipStream_t stream);

// Pair path: out per (first[i], second[i]); invalid ids give NaN.
hipError_t run_pairs(int mode, bool weighted, int32_t n, const int32_t* off, const int32_t* ind,
                     const float* vw, const float* work32, int64_t np, const int32_t* first,
                     const int32_t* second, float* out32, float* ws32, int accumulate,
                     hipStream_t stream);

// Fills *info (device memory) for vertex weights vw[n] and their row sums S[n].
hipError_t launch_exact_info(int32_t n, const float* vw, const double* S, ExactInfo* info, hipStream_t s);

// Records a failing HIP call for jc_last_hip_error(); returns the JC_ERR_* code.
int note_hip_error(hipError_t e);

// Streaming kernels.
hipError_t launch_fill_f32(float* x, int64_t n, float v, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* x, int64_t n, uint32_t v, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* x, int64_t n, uint64_t v, hipStream_t s);
// Row sums take a scratch area of rowsum_scratch_bytes(n) (long-row list),
// or nullptr to allocate one on the stream for the call.
size_t rowsum_scratch_bytes(int32_t n);
hipError_t launch_rowsum_f32(int32_t n, const int32_t* off, const int32_t* ind, const float* v,
                             float* work, void* scratch, hipStream_t s);
hipError_t launch_rowsum_f64(int32_t n, const int32_t* off, const int32_t* ind, const float* v,
                             double* S, void* scratch, hipStream_t s);
hipError_t launch_vertex_weights(int32_t n, const int32_t* off, const float* edge_w, float* W,
                                 void* scratch, hipStream_t s);
hipError_t launch_jw_f32(int64_t m, const float* wi, const float* ws, float* out, hipStream_t s);
hipError_t launch_jw_f64(int64_t m, const double* wi, const double* ws, float* out, hipStream_t s);
hipError_t launch_ec_scan(int32_t n, const int32_t* off, int32_t* dest, float* wj, hipStream_t s);
// bitmap table region of the workgroup kernel with one workgroup per CU:
// [0] unweighted, [1] weighted layout
void bitmap_table_bytes(int64_t out[2]);
// jc_degree_stats: 69 words (see k_degree_stats), zeroed by the caller
hipError_t launch_degree_stats(int32_t n, const int32_t* off, const int32_t* ind, unsigned long long* out,
                               hipStream_t s);
hipError_t launch_count_above(int64_t m, const float* x, float thresh, unsigned long long* cnt,
                              hipStream_t s);
hipError_t launch_validate(int32_t n, int64_t m, const int32_t* off, const int32_t* ind,
                           int* flags, hipStream_t s);

int device_cu_count();

}  // namespace jc
