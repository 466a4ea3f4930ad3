// Owner engine: degree-binned sorted-list intersection for every stored edge.
//
// Reference counterparts: the per-edge loop of Jaccard_ec_unweighted
// (jaccard.cpp:711-742) and the per-row intersection of Jaccard_IsKernel
// (jaccard.cpp:329-398). Both compute, for a stored edge (u, v),
// I = |N(u) & N(v)| (or the vertex-weighted sum over the intersection) by
// binary searching the shorter row's columns in the longer row.
//
// Design for gfx950:
//   * Each undirected pair {u, x} is computed once, by the endpoint with the
//     larger (degree, id) -- its "owner" -- and written to both stored
//     directions (the mirror position is found while streaming or by a wave
//     search). Work is sum over pairs of min(d_u, d_x), half of the reference.
//   * The owner's row is staged in LDS as an open-addressing hash table (key
//     = column id; weighted: key + fp32 vertex weight in one 8-byte slot).
//     The short row is streamed from HBM with coalesced 256-byte wave loads
//     and each element is one LDS probe, instead of a dependent binary search
//     through HBM.
//   * Degree binning: rows with degree <= kSmallMax run one wave per row with a
//     per-wave table; longer rows run one workgroup per (row, 512-entry piece)
//     task from a degree-descending queue (persistent grid, atomic counter).
//     Rows longer than the LDS chunk are split into chunks whose tables are
//     built one after the other; per-item cursors/partials live in LDS.
//   * Weighted sums are double-double. For fp32 addends whose bit span fits
//     in 106 bits the sum is exact, so results do not depend on chunking,
//     task split or lane order: plain, batched and pair paths agree bitwise.
//   * Positions that no symmetric pair reaches (asymmetric CSR input) keep a
//     sentinel and are recomputed by a fix-up pass with the pair routine.
#include "jc_device.hpp"
#include "jc_internal.hpp"
#include "jc_profile.hpp"

namespace jc {

struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (b.hi - bb);
  double t = a.lo + b.lo;
  t += e;
  const double hi = s + t;
  return DD{hi, t - (hi - s)};
}

__device__ __forceinline__ DD dd_add1(DD a, double v) {
  const double s = a.hi + v;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (v - bb);
  const double t = a.lo + e;
  const double hi = s + t;
  return DD{hi, t - (hi - s)};
}

__device__ __forceinline__ DD wave_sum_dd(DD v) {
#pragma unroll
  for (int s = 1; s < kWave; s <<= 1) {
    DD o{__shfl_xor(v.hi, s), __shfl_xor(v.lo, s)};
    v = dd_add(v, o);
  }
  return v;
}

__device__ __forceinline__ void lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

// ---------------------------------------------------------------------------
// LDS hash tables.
template <bool W>
struct Table;

template <>
struct Table<false> {
  using Slot = int32_t;
  Slot* s;
  int shift;
  uint32_t mask;
  __device__ __forceinline__ void clear(int i) const { s[i] = -1; }
  __device__ __forceinline__ void insert(int32_t y, float) const {
    uint32_t h = hslot(y, shift);
    for (;;) {
      const int32_t old = atomicCAS(&s[h], -1, y);
      if (old == -1 || old == y) return;
      h = (h + 1) & mask;
    }
  }
  __device__ __forceinline__ bool probe(int32_t y, float&) const {
    uint32_t h = hslot(y, shift);
    for (;;) {
      const int32_t k = s[h];
      if (k == y) return true;
      if (k < 0) return false;
      h = (h + 1) & mask;
    }
  }
};

template <>
struct Table<true> {
  using Slot = int2;  // .x key, .y fp32 weight bits
  Slot* s;
  int shift;
  uint32_t mask;
  __device__ __forceinline__ void clear(int i) const { s[i] = make_int2(-1, 0); }
  __device__ __forceinline__ void insert(int32_t y, float v) const {
    uint32_t h = hslot(y, shift);
    for (;;) {
      const int32_t old = atomicCAS(&s[h].x, -1, y);
      if (old == -1) {
        s[h].y = __float_as_int(v);
        return;
      }
      if (old == y) return;
      h = (h + 1) & mask;
    }
  }
  __device__ __forceinline__ bool probe(int32_t y, float& v) const {
    uint32_t h = hslot(y, shift);
    for (;;) {
      const int2 k = s[h];
      if (k.x == y) {
        v = __int_as_float(k.y);
        return true;
      }
      if (k.x < 0) return false;
      h = (h + 1) & mask;
    }
  }
};

// ---------------------------------------------------------------------------
// Epilogue: write the result of pair (u, x) at position k and, if q >= 0, at
// the mirror position q. Called by one lane.
template <int MODE, bool W>
__device__ __forceinline__ void emit(const EngineArgs& a, int32_t u, int du, int32_t x, int dx,
                                     int64_t k, int64_t q, uint32_t cnt, DD sum) {
  if constexpr (MODE == kModeEcU) {
    const float fI = (float)cnt;
    const float r = fI / ((float)(du + dx) - fI);
    a.out32[k] = r;
    if (q >= 0) a.out32[q] = r;
  } else if constexpr (MODE == kModeEcW) {
    const double ws = a.S64[u] + a.S64[x];
    const float r = (float)(sum.hi / (ws - sum.hi));
    a.out32[k] = r;
    if (q >= 0) a.out32[q] = r;
  } else if constexpr (MODE == kModeIsF32) {
    const float add = W ? (float)sum.hi : (float)cnt;
    const float ws = a.work32[u] + a.work32[x];
    a.out32[k] = (a.accumulate ? a.out32[k] : 0.0f) + add;
    a.ws32[k] = ws;
    if (q >= 0) {
      a.out32[q] = (a.accumulate ? a.out32[q] : 0.0f) + add;
      a.ws32[q] = ws;
    }
  } else {
    const double ws = a.S64[u] + a.S64[x];
    a.wi64[k] = sum.hi;
    a.ws64[k] = ws;
    if (q >= 0) {
      a.wi64[q] = sum.hi;
      a.ws64[q] = ws;
    }
  }
}

// Streams the whole short row [ox, ox+dx), probes elements inside [lo, hi],
// and locates u inside it (the mirror position). Wave-uniform results.
template <bool W>
__device__ __forceinline__ void full_pass(const int32_t* __restrict__ ind, int ox, int dx, int32_t u,
                                          int32_t lo, int32_t hi, const Table<W>& tab,
                                          uint32_t& cnt, DD& acc, int& below, bool& found) {
  const int end = ox + dx;
  for (int p0 = ox; p0 < end; p0 += kWave) {
    const int p = p0 + lane_id();
    const bool act = p < end;
    const int32_t y = act ? ind[p] : INT32_MAX;
    below += popc64(ballot(act && y < u));
    found |= ballot(act && y == u) != 0;
    bool hit = false;
    float v = 0.0f;
    if (act && y >= lo && y <= hi) hit = tab.probe(y, v);
    cnt += popc64(ballot(hit));
    if constexpr (W) {
      if (hit) acc = dd_add1(acc, (double)v);
    }
  }
}

// Streams positions [s, e) of a short row and probes every element.
template <bool W>
__device__ __forceinline__ void range_pass(const int32_t* __restrict__ ind, int s, int e, int32_t lo,
                                           int32_t hi, const Table<W>& tab, uint32_t& cnt, DD& acc) {
  for (int p0 = s; p0 < e; p0 += kWave) {
    const int p = p0 + lane_id();
    bool hit = false;
    float v = 0.0f;
    if (p < e) {
      const int32_t y = ind[p];
      if (y >= lo && y <= hi) hit = tab.probe(y, v);
    }
    cnt += popc64(ballot(hit));
    if constexpr (W) {
      if (hit) acc = dd_add1(acc, (double)v);
    }
  }
}

// ---------------------------------------------------------------------------
// Rows with degree <= kSmallMax: one wave per row, a private table per wave.
template <int MODE, bool W, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_owner_small(EngineArgs a) {
  constexpr int NW = BLOCK / kWave;
  constexpr int PW = 2 * kSmallMax;
  using Slot = typename Table<W>::Slot;
  __shared__ Slot tab_all[NW][PW];
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int nsmall = *a.num_small;
  const int stride = gridDim.x * NW;
  for (int r = blockIdx.x * NW + w; r < nsmall; r += stride) {
    const int32_t u = a.small_rows[r];
    const int ou = a.off[u];
    const int du = a.off[u + 1] - ou;
    // owned neighbours of u (du <= 128: two lane groups)
    int32_t x0 = 0, x1 = 0;
    int dx0 = 0, dx1 = 0, ox0 = 0, ox1 = 0;
    if (lane < du) {
      x0 = a.ind[ou + lane];
      ox0 = a.off[x0];
      dx0 = a.off[x0 + 1] - ox0;
    }
    if (lane + kWave < du) {
      x1 = a.ind[ou + kWave + lane];
      ox1 = a.off[x1];
      dx1 = a.off[x1 + 1] - ox1;
    }
    uint64_t own0 = ballot(lane < du && owns(u, du, x0, dx0));
    uint64_t own1 = ballot(lane + kWave < du && owns(u, du, x1, dx1));
    if ((own0 | own1) == 0) continue;
    const int logP = max(ceil_log2(2 * du), 5);
    const Table<W> tab{tab_all[w], 32 - logP, (uint32_t)((1 << logP) - 1)};
    for (int i = lane; i < (1 << logP); i += kWave) tab.clear(i);
    lds_fence();
    for (int i = lane; i < du; i += kWave) {
      const int32_t y = a.ind[ou + i];
      tab.insert(y, W ? a.vw[y] : 0.0f);
    }
    lds_fence();
    const int32_t lo = a.ind[ou], hi = a.ind[ou + du - 1];
    for (int g = 0; g < 2; ++g) {
      uint64_t own = g == 0 ? own0 : own1;
      while (own) {
        const int l = __ffsll((unsigned long long)own) - 1;
        own &= own - 1;
        const int32_t x = __shfl(g == 0 ? x0 : x1, l);
        const int ox = __shfl(g == 0 ? ox0 : ox1, l);
        const int dx = __shfl(g == 0 ? dx0 : dx1, l);
        uint32_t cnt = 0;
        DD acc{0.0, 0.0};
        int below = 0;
        bool found = false;
        full_pass<W>(a.ind, ox, dx, u, lo, hi, tab, cnt, acc, below, found);
        if constexpr (W) acc = wave_sum_dd(acc);
        if (lane == 0)
          emit<MODE, W>(a, u, du, x, dx, (int64_t)ou + g * kWave + l, found ? (int64_t)ox + below : -1,
                        cnt, acc);
      }
    }
    lds_fence();  // all probes of this row done before the next row clears the table
  }
}

// ---------------------------------------------------------------------------
// Rows with degree > kSmallMax: one workgroup per (row, piece) task.
template <bool W>
struct BlockCfg {
  static constexpr int CH = W ? kChunkW : kChunkU;
  static constexpr int PMAX = 2 * CH;
  using Slot = typename Table<W>::Slot;
  static constexpr size_t table_bytes = (size_t)PMAX * sizeof(Slot);
  static constexpr size_t acc_bytes = (size_t)kTaskSpan * (W ? sizeof(DD) : sizeof(uint32_t));
  static constexpr size_t cursor_bytes = (size_t)kTaskSpan * sizeof(int);
  static constexpr size_t lds_bytes = table_bytes + acc_bytes + cursor_bytes;
};

// Positions a wave takes from the task per grab: small grabs keep the waves of
// a workgroup balanced when neighbour degrees vary widely.
constexpr int kGrab = 16;

template <int MODE, bool W, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_owner_block(EngineArgs a) {
  using Cfg = BlockCfg<W>;
  using Slot = typename Cfg::Slot;
  constexpr int CH = Cfg::CH;
  extern __shared__ __align__(16) unsigned char smem[];
  Slot* tslots = reinterpret_cast<Slot*>(smem);
  DD* acc_dd = reinterpret_cast<DD*>(smem + Cfg::table_bytes);
  uint32_t* acc_cnt = reinterpret_cast<uint32_t*>(smem + Cfg::table_bytes);
  int* cursor = reinterpret_cast<int*>(smem + Cfg::table_bytes + Cfg::acc_bytes);
  __shared__ int s_task;
  __shared__ int s_next;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int ntasks = *a.num_tasks;
  for (;;) {
    __syncthreads();
    if (tid == 0) s_task = atomicAdd(a.task_counter, 1);
    __syncthreads();
    const int t = s_task;
    if (t >= ntasks) return;
    const int2 tk = a.tasks[t];
    const int32_t u = tk.x;
    const int ou = a.off[u];
    const int du = a.off[u + 1] - ou;
    const int ka = ou + tk.y * kTaskSpan;
    const int span = min(kTaskSpan, ou + du - ka);
    const int nch = (du + CH - 1) / CH;
    for (int c = 0; c < nch; ++c) {
      const int cs = ou + c * CH;
      const int ce = min(cs + CH, ou + du);
      const int clen = ce - cs;
      const int logP = max(ceil_log2(2 * clen), 5);
      const Table<W> tab{tslots, 32 - logP, (uint32_t)((1 << logP) - 1)};
      for (int i = tid; i < (1 << logP); i += BLOCK) tab.clear(i);
      if (tid == 0) s_next = 0;
      __syncthreads();
      for (int i = tid; i < clen; i += BLOCK) {
        const int32_t y = a.ind[cs + i];
        tab.insert(y, W ? a.vw[y] : 0.0f);
      }
      __syncthreads();
      const int32_t lo = a.ind[cs], hi = a.ind[ce - 1];
      const bool last = c == nch - 1;
      for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&s_next, kGrab);
        base = __shfl(base, 0);
        if (base >= span) break;
        const int j = base + lane;
        int32_t x = 0;
        int ox = 0, dx = 0;
        if (lane < kGrab && j < span) {
          x = a.ind[ka + j];
          ox = a.off[x];
          dx = a.off[x + 1] - ox;
        }
        uint64_t own = ballot(lane < kGrab && j < span && owns(u, du, x, dx));
        while (own) {
          const int l = __ffsll((unsigned long long)own) - 1;
          own &= own - 1;
          const int jj = base + l;
          const int32_t xx = __shfl(x, l);
          const int oxx = __shfl(ox, l);
          const int dxx = __shfl(dx, l);
          uint32_t cnt = 0;
          DD acc{0.0, 0.0};
          if (nch == 1) {
            int below = 0;
            bool found = false;
            full_pass<W>(a.ind, oxx, dxx, u, lo, hi, tab, cnt, acc, below, found);
            if constexpr (W) acc = wave_sum_dd(acc);
            if (lane == 0)
              emit<MODE, W>(a, u, du, xx, dxx, (int64_t)ka + jj, found ? (int64_t)oxx + below : -1,
                            cnt, acc);
            continue;
          }
          const int end = oxx + dxx;
          const int s = c == 0 ? oxx : cursor[jj];
          int e = s;
          if (s < end) {
            if (last)
              e = end;
            else if (a.ind[s] <= hi)
              e = wave_lower_bound(a.ind, s, end, hi + 1);
          }
          range_pass<W>(a.ind, s, e, lo, hi, tab, cnt, acc);
          if constexpr (W) acc = wave_sum_dd(acc);
          if (!last) {
            if (lane == 0) {
              cursor[jj] = e;
              if constexpr (W) {
                acc_dd[jj] = c == 0 ? acc : dd_add(acc_dd[jj], acc);
              } else {
                acc_cnt[jj] = c == 0 ? cnt : acc_cnt[jj] + cnt;
              }
            }
          } else {
            if constexpr (W) {
              acc = dd_add(acc_dd[jj], acc);
            } else {
              cnt += acc_cnt[jj];
            }
            const int q = wave_lower_bound(a.ind, oxx, end, u);
            const bool found = q < end && a.ind[q] == u;
            if (lane == 0) emit<MODE, W>(a, u, du, xx, dxx, (int64_t)ka + jj, found ? (int64_t)q : -1, cnt, acc);
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Generic pair routine (one wave): intersection of N(a) and N(b) by per-lane
// binary search of the shorter row's elements in the longer row. Used by the
// fix-up pass and the pair-list path.
template <bool W>
__device__ __forceinline__ void wave_pair(const int32_t* __restrict__ off, const int32_t* __restrict__ ind,
                                          const float* __restrict__ vw, int32_t pa, int32_t pb,
                                          uint32_t& cnt, DD& acc) {
  const int oa = off[pa], da = off[pa + 1] - oa;
  const int ob = off[pb], db = off[pb + 1] - ob;
  const bool a_long = da > db || (da == db && pa >= pb);
  const int oL = a_long ? oa : ob, dL = a_long ? da : db;
  const int ox = a_long ? ob : oa, dx = a_long ? db : da;
  cnt = 0;
  acc = DD{0.0, 0.0};
  if (dx == 0) return;
  const int32_t lo = ind[oL], hi = ind[oL + dL - 1];
  for (int p0 = ox; p0 < ox + dx; p0 += kWave) {
    const int p = p0 + lane_id();
    bool hit = false;
    int32_t y = 0;
    if (p < ox + dx) {
      y = ind[p];
      if (y >= lo && y <= hi) {
        const int pos = lane_lower_bound(ind, oL, oL + dL, y);
        hit = pos < oL + dL && ind[pos] == y;
      }
    }
    cnt += popc64(ballot(hit));
    if constexpr (W) {
      if (hit) acc = dd_add1(acc, (double)vw[y]);
    }
  }
  if constexpr (W) acc = wave_sum_dd(acc);
}

// Exact row sum of vw over N(r), one wave.
__device__ __forceinline__ double wave_row_wsum(const int32_t* __restrict__ off,
                                                const int32_t* __restrict__ ind,
                                                const float* __restrict__ vw, int32_t r) {
  DD acc{0.0, 0.0};
  const int o = off[r], d = off[r + 1] - o;
  for (int p0 = o; p0 < o + d; p0 += kWave) {
    const int p = p0 + lane_id();
    if (p < o + d) acc = dd_add1(acc, (double)vw[ind[p]]);
  }
  return wave_sum_dd(acc).hi;
}

template <bool W64>
__device__ __forceinline__ bool is_sentinel(const void* arr, int64_t k) {
  if constexpr (W64)
    return reinterpret_cast<const uint64_t*>(arr)[k] == kSentinel64;
  else
    return reinterpret_cast<const uint32_t*>(arr)[k] == kSentinel32;
}

// Fix-up: recompute positions still holding the sentinel.
template <int MODE, bool W, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_fixup(EngineArgs a) {
  constexpr bool F64 = MODE == kModeIsF64;
  const void* flag_arr = MODE == kModeEcU || MODE == kModeEcW ? (const void*)a.out32
                         : MODE == kModeIsF32                ? (const void*)a.ws32
                                                             : (const void*)a.ws64;
  const int64_t nwaves = (int64_t)gridDim.x * (BLOCK / kWave);
  const int64_t gw = (int64_t)blockIdx.x * (BLOCK / kWave) + threadIdx.x / kWave;
  for (int64_t base = gw * kWave; base < a.m; base += nwaves * kWave) {
    const int64_t k = base + lane_id();
    uint64_t bad = ballot(k < a.m && is_sentinel<F64>(flag_arr, k));
    while (bad) {
      const int l = __ffsll((unsigned long long)bad) - 1;
      bad &= bad - 1;
      const int64_t kk = base + l;
      const int32_t u = wave_lower_bound(a.off, 0, a.n + 1, (int32_t)(kk + 1)) - 1;
      const int32_t v = a.ind[kk];
      uint32_t cnt;
      DD acc;
      wave_pair<W>(a.off, a.ind, a.vw, u, v, cnt, acc);
      if (lane_id() == 0) {
        const int du = a.off[u + 1] - a.off[u];
        const int dv = a.off[v + 1] - a.off[v];
        emit<MODE, W>(a, u, du, v, dv, kk, -1, cnt, acc);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Schedule: degree-binned task list, built on the device without a host sync.
// Buckets 0..31 by floor(log2(degree)) for rows with degree > kSmallMax; the
// task list is laid out from the highest bucket down so the persistent
// workgroups take the heaviest rows first.
constexpr int kBuckets = 32;

struct Sched {
  int bucket_cnt[kBuckets];   // tasks per bucket
  int bucket_cur[kBuckets];   // emission cursors
  int num_tasks;
  int task_counter;
  int num_small;
  int small_cur;
};

__device__ __forceinline__ int bucket_of(int d) { return 31 - __clz(d); }

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sched_count(int32_t n, const int32_t* __restrict__ off,
                                                       Sched* sc) {
  __shared__ int hist[kBuckets + 1];
  for (int i = threadIdx.x; i < kBuckets + 1; i += BLOCK) hist[i] = 0;
  __syncthreads();
  for (int32_t u = blockIdx.x * BLOCK + threadIdx.x; u < n; u += gridDim.x * BLOCK) {
    const int d = off[u + 1] - off[u];
    if (d <= 0) continue;
    if (d <= kSmallMax)
      atomicAdd(&hist[kBuckets], 1);
    else
      atomicAdd(&hist[bucket_of(d)], (d + kTaskSpan - 1) / kTaskSpan);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBuckets; i += BLOCK)
    if (hist[i]) atomicAdd(&sc->bucket_cnt[i], hist[i]);
  if (threadIdx.x == 0 && hist[kBuckets]) atomicAdd(&sc->num_small, hist[kBuckets]);
}

__global__ void k_sched_init(Sched* sc) {
  int* p = reinterpret_cast<int*>(sc);
  for (int i = threadIdx.x; i < (int)(sizeof(Sched) / sizeof(int)); i += blockDim.x) p[i] = 0;
}

__global__ void k_sched_scan(Sched* sc) {
  if (threadIdx.x != 0) return;
  int acc = 0;
  for (int b = kBuckets - 1; b >= 0; --b) {
    sc->bucket_cur[b] = acc;
    acc += sc->bucket_cnt[b];
  }
  sc->num_tasks = acc;
  sc->task_counter = 0;
  sc->small_cur = 0;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sched_emit(int32_t n, const int32_t* __restrict__ off,
                                                      Sched* sc, int2* tasks, int32_t* small_rows) {
  __shared__ int hist[kBuckets + 1];
  __shared__ int basep[kBuckets + 1];
  for (int32_t u0 = blockIdx.x * BLOCK; u0 < n; u0 += gridDim.x * BLOCK) {
    for (int i = threadIdx.x; i < kBuckets + 1; i += BLOCK) hist[i] = 0;
    __syncthreads();
    const int32_t u = u0 + threadIdx.x;
    int d = 0, b = -1, cntv = 0, local = 0;
    if (u < n) {
      d = off[u + 1] - off[u];
      if (d > 0) {
        if (d <= kSmallMax) {
          b = kBuckets;
          cntv = 1;
        } else {
          b = bucket_of(d);
          cntv = (d + kTaskSpan - 1) / kTaskSpan;
        }
        local = atomicAdd(&hist[b], cntv);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kBuckets + 1; i += BLOCK) {
      if (hist[i]) basep[i] = i == kBuckets ? atomicAdd(&sc->small_cur, hist[i]) : atomicAdd(&sc->bucket_cur[i], hist[i]);
    }
    __syncthreads();
    if (b == kBuckets) {
      small_rows[basep[b] + local] = u;
    } else if (b >= 0) {
      const int at = basep[b] + local;
      for (int p = 0; p < cntv; ++p) tasks[at + p] = make_int2(u, p);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host side.
static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static int64_t max_tasks(int32_t n, int64_t m) {
  // each row with degree > kSmallMax contributes ceil(d / kTaskSpan) <= d / kTaskSpan + 1 tasks
  return m / kTaskSpan + std::min<int64_t>(n, m / (kSmallMax + 1)) + 1;
}

size_t engine_workspace_bytes(int32_t n, int64_t m) {
  size_t b = align_up(sizeof(Sched), 256);
  b += align_up((size_t)max_tasks(n, m) * sizeof(int2), 256);
  b += align_up((size_t)std::max<int32_t>(n, 1) * sizeof(int32_t), 256);
  return b;
}

template <int MODE, bool W>
static hipError_t run_engine_t(const EngineArgs& base, void* ws, bool fixup, hipStream_t stream) {
  constexpr int SB = 256;   // small-row kernel block
  constexpr int BB = 512;   // block kernel block
  constexpr int PB = 256;   // schedule kernels
  EngineArgs a = base;
  unsigned char* p = static_cast<unsigned char*>(ws);
  Sched* sc = reinterpret_cast<Sched*>(p);
  p += align_up(sizeof(Sched), 256);
  int2* tasks = reinterpret_cast<int2*>(p);
  p += align_up((size_t)max_tasks(a.n, a.m) * sizeof(int2), 256);
  int32_t* small = reinterpret_cast<int32_t*>(p);
  a.tasks = tasks;
  a.num_tasks = &sc->num_tasks;
  a.task_counter = &sc->task_counter;
  a.small_rows = small;
  a.num_small = &sc->num_small;

  const int cus = device_cu_count();
  hipError_t err;
  hipLaunchKernelGGL(k_sched_init, dim3(1), dim3(128), 0, stream, sc);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  if (a.n > 0) {
    const int g = std::min<int64_t>((a.n + PB - 1) / PB, (int64_t)cus * 8);
    {
      ProfScope ps("sched", stream);
      hipLaunchKernelGGL(k_sched_count<PB>, dim3(g), dim3(PB), 0, stream, a.n, a.off, sc);
      hipLaunchKernelGGL(k_sched_scan, dim3(1), dim3(64), 0, stream, sc);
      hipLaunchKernelGGL(k_sched_emit<PB>, dim3(g), dim3(PB), 0, stream, a.n, a.off, sc, tasks, small);
      if ((err = hipGetLastError()) != hipSuccess) return err;
    }

    using Cfg = BlockCfg<W>;
    err = hipFuncSetAttribute(reinterpret_cast<const void*>(k_owner_block<MODE, W, BB>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)Cfg::lds_bytes);
    if (err != hipSuccess) return err;
    int occ = 0;
    err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_owner_block<MODE, W, BB>, BB, Cfg::lds_bytes);
    if (err != hipSuccess) return err;
    if (occ < 1) return hipErrorInvalidConfiguration;
    {
      ProfScope ps("owner_block", stream);
      hipLaunchKernelGGL((k_owner_block<MODE, W, BB>), dim3(cus * occ), dim3(BB), Cfg::lds_bytes, stream, a);
      if ((err = hipGetLastError()) != hipSuccess) return err;
    }
    const int sg = std::min<int64_t>((a.n + (SB / kWave) - 1) / (SB / kWave), (int64_t)cus * 16);
    {
      ProfScope ps("owner_small", stream);
      hipLaunchKernelGGL((k_owner_small<MODE, W, SB>), dim3(sg), dim3(SB), 0, stream, a);
      if ((err = hipGetLastError()) != hipSuccess) return err;
    }
  }
  if (fixup && a.m > 0) {
    ProfScope ps("fixup", stream);
    const int fg = std::min<int64_t>((a.m + 255) / 256, (int64_t)cus * 8);
    hipLaunchKernelGGL((k_fixup<MODE, W, 256>), dim3(fg), dim3(256), 0, stream, a);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  }
  return hipSuccess;
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
