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
//     directions (the mirror position is counted while streaming, or found
//     by one group search for rows split over several chunks). Work is the
//     sum over pairs of min(d_u, d_x): half of the reference's.
//   * The owner's row is staged in LDS as a bucketized hash table: buckets
//     of 8 slots (32 bytes as two 16-byte halves in separate arrays, two
//     ds_read_b128), at most two keys per bucket on average weighted and
//     three unweighted, weights in a parallel array; rows whose id span fits
//     use an LDS bitmap instead. A probe is straight-line code; a full
//     bucket (rare) sends the wave through a uniform slow path. The short
//     row is streamed from HBM with raw buffer loads, several in flight per
//     lane.
//   * A wave is split into groups of G lanes, one short row per group, G and
//     the unroll U chosen by the row's length class, so short rows do not
//     leave lanes idle. Counts accumulate per lane (VALU) and are reduced
//     once per row with DPP lane moves; the next pass's loads are issued
//     before the current pass is probed.
//   * Degree binning: rows with degree <= kSmallMax run one wave per row with
//     a per-wave table; longer rows run one workgroup per (row, 896-entry
//     piece) task from degree-descending queues (persistent grids, atomic
//     counters), one queue and kernel instantiation per table kind (LDS
//     bitmap when the row's id span fits, else hash table). A task sorts its
//     owned entries into length classes, heaviest first, and its waves take
//     class units from an LDS counter. Bitmap tasks look one task ahead: wave
//     0 stages the next task's entries in LDS while the others run units.
//   * Rows longer than one LDS chunk are processed chunk by chunk; each short
//     row resumes where the previous chunk stopped (no searches), its
//     partial result kept in LDS.
//   * Weighted sums are double-double. For fp32 addends whose bit span (from
//     the finest set bit of any addend to the top of the largest sum) is at
//     most 104 bits every add is exact (each step rounds only sums of two
//     error terms, which need at most span - 51 bits), so the sum is exact, so results do not depend on chunking,
//     task split or lane order: plain, batched and pair paths agree bitwise.
//     When the weights' range makes plain fp64 sums exact (ExactInfo, checked
//     per call on the device) the kernels add in fp64, with the same result.
//   * Positions that no symmetric pair reaches (asymmetric CSR input) keep a
//     sentinel and are recomputed by a fix-up pass with the pair routine.
#pragma once
#include "jc_device.hpp"
#include "jc_hash.hpp"
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

// True when the range facts of this pass (ExactInfo) show every weight sum is
// exact in plain fp64 (exact_sums_fit_f64); without them, double-double.
__device__ __forceinline__ bool exact_f64(const EngineArgs& a) {
  if (a.exact == nullptr) return false;
  return exact_sums_fit_f64(*a.exact);
}

// ---------------------------------------------------------------------------
// Bucketized LDS hash table: 2^lb buckets of kBucket int32 slots. Slots 0..6
// hold keys (-1 = empty); slot 7 is the spill marker, set when a key had to
// move on to the next bucket. The slots are stored as two arrays of 16-byte
// halves: slots 0..3 of bucket b at key[4b], slots 4..7 at key[hofs + 4b]
// (hofs = 4 * the table's bucket capacity, a constant of the kernel, so the
// second read's offset is an immediate). Consecutive buckets' halves are
// consecutive 16-byte words: the two ds_read_b128 of a probe each spread
// over every bank group of the LDS, where adjacent 32-byte buckets would put
// each read on half of them. Weighted tables keep the fp32 weight of key
// slot i of bucket b at val[8b + i]; val[7] (bucket 0's unused slot 7) holds
// 0.0f and is what a missing lane reads, so every lane issues the same one
// weight read without a branch and the missing lanes share one LDS word
// (a broadcast, no bank conflicts). Bucket = top bits of a multiplicative
// hash.
#ifndef JC_F64FAST
#define JC_F64FAST 1
#endif
#ifndef JC_FETCHALL
#define JC_FETCHALL 1
#endif
#ifndef JC_MIRRORCOUNT
#define JC_MIRRORCOUNT 1
#endif
#ifndef JC_PINGPONG
#define JC_PINGPONG 1  // row streaming loop unrolled twice (two register sets, no copies)
#endif
constexpr int kBucket = 8;
constexpr int kBucketKeys = 7;
constexpr int32_t kSpill = -2;

struct Hit {
  bool hit;
  bool spill;
  float v;
};

// Table kinds of the owner engine's LDS tables.
constexpr int kTkHash = 0;    // 32-bit keys (Table)
constexpr int kTkBitmap = 1;  // bitmap over the row's id span
constexpr int kTkQuot = 2;    // 16-bit quotient keys (QTable)

template <bool W>
struct Table {
  int32_t* key;
  float* val;
  int lb;           // log2 of the bucket count
  uint32_t bmask;   // buckets - 1
  int hofs;         // ints from the first half array to the second
  __device__ __forceinline__ uint32_t bucket(int32_t y) const { return hash_bucket(y, lb); }
  __device__ __forceinline__ uint32_t next(uint32_t b) const { return (b + 1) & bmask; }
  __device__ __forceinline__ int32_t* slot_ptr(uint32_t b, int s) const {
    return key + (s < 4 ? 0 : hofs) + 4 * b + (s & 3);
  }
  // clears the i-th 4-slot word of both halves, 0 <= i < 4 << lb
  __device__ __forceinline__ void clear(int i) const {
    key[i] = -1;
    key[hofs + i] = -1;
  }
  // the miss word; written once per table build, before any probe
  __device__ __forceinline__ void set_zero() const {
    if constexpr (W) val[kBucketKeys] = 0.0f;
  }
  __device__ __forceinline__ void insert(int32_t y, float v) const {
    uint32_t b = bucket(y);
    for (;;) {
      for (int s = 0; s < kBucketKeys; ++s) {
        const int32_t old = atomicCAS(slot_ptr(b, s), -1, y);
        if (old == -1) {
          if constexpr (W) val[b * kBucket + s] = v;
          return;
        }
        if (old == y) return;
      }
      *slot_ptr(b, kBucketKeys) = kSpill;
      b = next(b);
    }
  }
  // Straight-line probe of bucket b: two ds_read_b128 for the keys, seven
  // compares; weighted tables add one ds_read_b32 at the hit slot (slot
  // index from three masks: bit 0 = slots 1, 3, 5; bit 1 = 2, 3, 6; bit 2 =
  // 4, 5, 6) or at the miss word.
  __device__ __forceinline__ void fetch(uint32_t b, int4& k0, int4& k1) const {
    k0 = *reinterpret_cast<const int4*>(key + 4 * b);
    k1 = *reinterpret_cast<const int4*>(key + hofs + 4 * b);
  }
  __device__ __forceinline__ Hit look(uint32_t b, int32_t y) const {
    int4 k0, k1;
    fetch(b, k0, k1);
    return match(b, y, k0, k1);
  }
  // the compares and the weight read of look, on keys already fetched
  __device__ __forceinline__ Hit match(uint32_t b, int32_t y, const int4& k0, const int4& k1) const {
    const bool e0 = k0.x == y, e1 = k0.y == y, e2 = k0.z == y, e3 = k0.w == y;
    const bool e4 = k1.x == y, e5 = k1.y == y, e6 = k1.z == y;
    Hit h;
    h.hit = e0 | e1 | e2 | e3 | e4 | e5 | e6;
    h.spill = k1.w == kSpill;
    h.v = 0.0f;
    if constexpr (W) {
      const bool m1 = e1 | e3 | e5, m2 = e2 | e3 | e6, m4 = e4 | e5 | e6;
      const uint32_t slot = (m1 ? 1u : 0u) + (m2 ? 2u : 0u) + (m4 ? 4u : 0u);
      uint32_t idx = b * kBucket + slot;
      asm volatile("" : "+v"(idx));  // computed on every lane (see QTable::match)
      h.v = val[h.hit ? idx : (uint32_t)kBucketKeys];
    }
    return h;
  }
};

// Average keys per bucket (of kBucketKeys key slots) at which the workgroup
// kernel cuts a long row into chunks: 3, weighted or not (weighted rows were
// cut at 2; fewer, fuller chunks measured faster, as chunk restarts cost more
// than the overflow walks they save: RMAT-22 weighted on 32-bit tables 68.5
// -> 65.8 ms). The one-wave-per-row kernel sizes its private tables for at
// most kLambdaSmall keys per bucket (2 weighted, where 3 measured slower).
#ifndef JC_LAMBDA_U
#define JC_LAMBDA_U 3
#endif
#ifndef JC_LAMBDA_W
#define JC_LAMBDA_W 3
#endif
#ifndef JC_LAMBDA_SMALL_W
#define JC_LAMBDA_SMALL_W 2
#endif
template <bool W>
constexpr int kLambda = W ? JC_LAMBDA_W : JC_LAMBDA_U;
template <bool W>
constexpr int kLambdaSmall = W ? JC_LAMBDA_SMALL_W : JC_LAMBDA_U;

// log2 of the bucket count for a chunk of `len` keys in the one-wave-per-row
// kernel: at most kLambdaSmall keys per bucket on average, at least 4 buckets.
template <bool W>
__device__ __forceinline__ int table_log_buckets(int len) {
  return max(ceil_log2((len + kLambdaSmall<W> - 1) / kLambdaSmall<W>), 2);
}

// The same for the workgroup kernel, whose table region always holds max_lb
// buckets: tables of chunks shorter than a full one get up to one bucket per
// key (fewer buckets overflow into their neighbours, which costs the whole
// wave an overflow walk on every pass where one lane meets one).
#ifndef JC_TABLE_LOWFILL
#define JC_TABLE_LOWFILL 1
#endif
template <bool W>
__device__ __forceinline__ int block_table_log_buckets(int len, int max_lb) {
  return JC_TABLE_LOWFILL ? min(max(ceil_log2(len), 2), max_lb) : table_log_buckets<W>(len);
}

// Inserts keys ind[cs .. cs+len) (weighted: with vw[key]) into the table,
// `stride` threads cooperating; four keys per thread are loaded before any
// insert so the loads overlap.
template <bool W>
__device__ __forceinline__ void build_table(const Table<W>& tab, const int32_t* __restrict__ ind,
                                            const float* __restrict__ vw, int cs, int len, int tid, int stride) {
  for (int i0 = tid; i0 < len; i0 += 4 * stride) {
    int32_t y[4];
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; ++t) y[t] = ind[cs + min(i0 + t * stride, len - 1)];
    if constexpr (W) {
#pragma unroll
      for (int t = 0; t < 4; ++t) v[t] = vw[y[t]];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (i0 + t * stride < len) tab.insert(y[t], v[t]);
  }
}

// ---------------------------------------------------------------------------
// Quotient hash table, for graphs with fewer than 2^24 vertices (the
// hash-table task kernel). The 24-bit hash word h = (y * 0x9E3779) mod 2^24
// is a bijection of [0, 2^24) (odd multiplier), so a key is known from its
// home bucket (the top lb bits of h) and the rb = 24 - lb low bits of h: a
// slot stores those rb bits and, above them, the key's displacement d from
// its home bucket (a key whose home bucket was full lives d buckets on), 16
// bits per key instead of 32. A bucket is one 16-byte word: slots 0..6 are
// the halves of words 0..2 and the low half of word 3 (0xFFFF = empty; slots
// fill in order); the high half of word 3 holds the spill flag (bit 15: a
// key was placed beyond this bucket) and, weighted, the bucket's base in a
// dense weight array (bits 0..14): the key in slot s of bucket b has weight
// val[1 + base(b) + s], and val[0] = 0.0f is what a missing lane reads. With
// lb >= kQMinLb the stored value (at most 2^16 - 2^rb - 1 + 2^rb - 1 for the
// largest displacement kept) never equals 0xFFFF. One 16-byte LDS read
// probes a bucket (the 32-bit table needs two), and with weights stored
// densely a weighted chunk holds three times the keys of the 32-bit layout
// (12288 against 4096), so rows need a third of the chunks.
constexpr int kQMinLb = 10;
constexpr int kQLambda = 3;  // average keys per bucket at full chunk size
// Graphs with fewer than kQuotMaxN vertices use them (jc_internal.hpp).
#ifndef JC_QTABLE
#define JC_QTABLE 1
#endif

template <bool W>
struct QTable {
  uint32_t* word;  // 4 words per bucket
  float* val;      // weighted: val[0] = 0, then the dense weights
  int lb;          // log2 of the bucket count (kQMinLb <= lb <= 14)
  uint32_t bmask;  // buckets - 1
  int rb;          // 24 - lb: bits of the hash word below the bucket
  int maxd;        // largest displacement a slot can hold: 2^(16 - rb) - 2
  __device__ __forceinline__ void split(int32_t y, uint32_t& b, uint32_t& r) const {
    const uint32_t mh = mul_u24_hash((uint32_t)y);
    b = __builtin_amdgcn_ubfe(mh, rb, lb);
    r = __builtin_amdgcn_ubfe(mh, 0, rb);
  }
  // clears bucket i: slots empty, spill flag and base 0
  __device__ __forceinline__ void clear(int i) const {
    *reinterpret_cast<uint4*>(word + 4 * i) = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0x0000FFFFu);
  }
  __device__ __forceinline__ uint4 fetch(uint32_t b) const { return *reinterpret_cast<const uint4*>(word + 4 * b); }
  // Places key y; on success pos = bucket * 8 + slot. False when every bucket
  // within the largest displacement is full (the task is then redone with the
  // 32-bit table). Keys of one table are distinct (rows are duplicate-free).
  __device__ bool insert(int32_t y, uint32_t& pos) const {
    uint32_t b, r;
    split(y, b, r);
    for (int d = 0; d <= maxd; ++d) {
      const uint32_t key16 = r | ((uint32_t)d << rb);
      // the bucket in one 16-byte LDS read per attempt (volatile: re-read
      // after a lost compare-and-swap; the LDS address space, or the generic
      // pointer would make it a flat load)
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      typedef __attribute__((address_space(3))) const volatile u32x4 lds_bucket;
      const lds_bucket* wb = (const lds_bucket*)(word + 4 * b);
      for (;;) {
        const u32x4 w4 = *wb;
        int s = -1;
        uint32_t old = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t wv = w4[k];
          if (s < 0 && (wv & 0xFFFFu) == 0xFFFFu) {
            s = 2 * k;
            old = wv;
          } else if (s < 0 && k < 3 && (wv >> 16) == 0xFFFFu) {
            s = 2 * k + 1;
            old = wv;
          }
        }
        if (s < 0) break;  // the bucket's seven slots are taken
        const uint32_t nv = (s & 1) ? ((old & 0xFFFFu) | (key16 << 16)) : ((old & 0xFFFF0000u) | key16);
        if (atomicCAS(word + 4 * b + (s >> 1), old, nv) == old) {
          pos = (b << 3) | (uint32_t)s;
          return true;
        }
      }
      atomicOr(word + 4 * b + 3, 0x80000000u);
      b = (b + 1) & bmask;
    }
    return false;
  }
  // Compares key16 with the seven slots of a fetched bucket; weighted tables
  // read the weight at the hit slot (slot index from three masks as in the
  // 32-bit table) or val[0].
  __device__ __forceinline__ Hit match(uint32_t key16, const uint4& w) const {
    const bool e0 = (uint16_t)w.x == (uint16_t)key16, e1 = (w.x >> 16) == key16;
    const bool e2 = (uint16_t)w.y == (uint16_t)key16, e3 = (w.y >> 16) == key16;
    const bool e4 = (uint16_t)w.z == (uint16_t)key16, e5 = (w.z >> 16) == key16;
    const bool e6 = (uint16_t)w.w == (uint16_t)key16;
    Hit h;
    h.hit = e0 | e1 | e2 | e3 | e4 | e5 | e6;
    h.spill = (int32_t)w.w < 0;
    h.v = 0.0f;
    if constexpr (W) {
      const bool m1 = e1 | e3 | e5, m2 = e2 | e3 | e6, m4 = e4 | e5 | e6;
      const uint32_t slot = (m1 ? 1u : 0u) + (m2 ? 2u : 0u) + (m4 ? 4u : 0u);
      const uint32_t base = __builtin_amdgcn_ubfe(w.w, 16, 15);
      uint32_t idx = 1u + base + slot;
      // The index is computed on every lane before the select: without the
      // empty asm the compiler sinks it and the load into a branch (measured
      // 6% slower on RMAT-22 weighted). The select is then one v_cndmask on
      // the compares' lane mask; a ballot-based select went through a VGPR
      // and a compare, two more instructions per probe (60.1 -> 58.7 ms).
      asm volatile("" : "+v"(idx));
      h.v = val[h.hit ? idx : 0u];
    }
    return h;
  }
};

// log2 of the bucket count of a quotient table for a chunk of len keys, at
// most max_lb (the table region's buckets). A table is sized for a mean of
// at most JC_QFILL2 / 2 keys per bucket where the region allows it: the
// fuller the buckets, the more of them overflow into a neighbour, and a pass
// in which any lane's bucket has overflowed walks the overflow for the whole
// wave. (Chunks are cut at kQLambda keys per bucket of the whole region.)
#ifndef JC_QFILL2
#define JC_QFILL2 2
#endif
__device__ __forceinline__ int qtable_log_buckets(int len, int max_lb) {
  return min(max(ceil_log2((2 * len + JC_QFILL2 - 1) / JC_QFILL2), kQMinLb), max_lb);
}

// ---------------------------------------------------------------------------
// Epilogue: write the result of pair (u, x) at position k and, if q >= 0, at
// the mirror position q. Called by one lane.
template <int MODE, bool W>
__device__ __forceinline__ void emit(const EngineArgs& a, int32_t u, int du, int32_t x, int dx,
                                     int64_t k, int64_t q, uint32_t cnt, DD sum, double su, double sx) {
  // su, sx: the row sums of u and x (S64, or work32 for kModeIsF32), loaded
  // by the caller ahead of time
  (void)u;
  (void)x;
  if constexpr (MODE == kModeEcU) {
    const float fI = (float)cnt;
    const float r = fI / ((float)(du + dx) - fI);
    a.out32[k] = r;
    if (q >= 0) a.out32[q] = r;
  } else if constexpr (MODE == kModeEcW) {
    const double ws = su + sx;
    const float r = (float)(sum.hi / (ws - sum.hi));
    a.out32[k] = r;
    if (q >= 0) a.out32[q] = r;
  } else if constexpr (MODE == kModeIsF32) {
    const float add = W ? (float)sum.hi : (float)cnt;
    const float ws = (float)su + (float)sx;
    a.out32[k] = (a.accumulate ? a.out32[k] : 0.0f) + add;
    a.ws32[k] = ws;
    if (q >= 0) {
      a.out32[q] = (a.accumulate ? a.out32[q] : 0.0f) + add;
      a.ws32[q] = ws;
    }
  } else {
    const double ws = su + sx;
    a.wi64[k] = sum.hi;
    a.ws64[k] = ws;
    if (q >= 0) {
      a.wi64[q] = sum.hi;
      a.ws64[q] = ws;
    }
  }
}

// Row sum of vertex r used by the epilogue of MODE (0 when none is needed).
template <int MODE>
__device__ __forceinline__ double row_sum_of(const EngineArgs& a, int32_t r) {
  if constexpr (MODE == kModeEcW || MODE == kModeIsF64)
    return a.S64[r];
  else if constexpr (MODE == kModeIsF32)
    return (double)a.work32[r];
  else
    return 0.0;
}

// ---------------------------------------------------------------------------
// Group primitives. A wave is split into 64/G groups of G lanes; each group
// works on its own item. Ballots are masked to the group, loops are
// group-uniform (groups of one wave may run different trip counts).
template <int G>
__device__ __forceinline__ uint64_t group_mask() {
  if constexpr (G == kWave)
    return ~0ull;
  else
    return ((1ull << G) - 1) << (lane_id() & ~(G - 1));
}

// lower_bound within one group: smallest i in [lo, hi) with a[i] >= key.
template <int G>
__device__ __forceinline__ int grp_lower_bound(const int32_t* __restrict__ a, int lo, int hi, int32_t key) {
  const uint64_t gm = group_mask<G>();
  const int gl = lane_id() & (G - 1);
  while (hi - lo > G) {
    const int stride = (hi - lo + G - 1) / G;
    const int idx = lo + gl * stride;
    const bool lt = idx < hi && a[idx] < key;
    const int c = popc64(ballot(lt) & gm);
    if (c == 0) return lo;
    const int nlo = lo + (c - 1) * stride + 1;
    const int nhi = min(hi, lo + c * stride);
    lo = nlo;
    hi = nhi;
  }
  const int idx = lo + gl;
  const bool lt = idx < hi && a[idx] < key;
  return lo + popc64(ballot(lt) & gm);
}

// Group sums over G lanes (G = 8, 16, 32, 64), the total in every lane of the
// group. Butterfly steps within a row of 16 lanes use DPP lane moves (VALU):
// quad_perm [1,0,3,2] and [2,3,0,1], then row_half_mirror and row_mirror
// (each pairs every lane with one in the other half of its 8 or 16 lanes).
// The 16-lane step uses ds_swizzle (xor 16 inside 32 lanes), and the 32-lane
// step two lane reads. All lanes of the wave must be active.
#ifndef JC_DPP
#define JC_DPP 1
#endif
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}

template <int STEP>
__device__ __forceinline__ int xlane_i(int v) {
#if JC_DPP
  if constexpr (STEP == 1) return dpp_mov<0xB1>(v);        // quad_perm [1,0,3,2]
  else if constexpr (STEP == 2) return dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  else if constexpr (STEP == 4) return dpp_mov<0x141>(v);  // row_half_mirror
  else if constexpr (STEP == 8) return dpp_mov<0x140>(v);  // row_mirror
  else return __builtin_amdgcn_ds_swizzle(v, 0x401F);      // xor 16 within 32 lanes
#else
  return __shfl_xor(v, STEP);
#endif
}

template <int STEP>
__device__ __forceinline__ double xlane_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = xlane_i<STEP>((int)(b & 0xFFFFFFFFll));
  const int hi = xlane_i<STEP>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int G>
__device__ __forceinline__ int grp_sum_i(int v) {
  v += xlane_i<1>(v);
  v += xlane_i<2>(v);
  if constexpr (G >= 8) v += xlane_i<4>(v);
  if constexpr (G >= 16) v += xlane_i<8>(v);
  if constexpr (G >= 32) v += xlane_i<16>(v);
  if constexpr (G >= 64) v = __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 32);
  return v;
}

// Plain fp64 group sum (exact when every partial sum is exact).
template <int G>
__device__ __forceinline__ double grp_sum_f64(double v) {
  v += xlane_d<1>(v);
  v += xlane_d<2>(v);
  if constexpr (G >= 8) v += xlane_d<4>(v);
  if constexpr (G >= 16) v += xlane_d<8>(v);
  if constexpr (G >= 32) v += xlane_d<16>(v);
  if constexpr (G >= 64) {
    const long long b = __double_as_longlong(v);
    const int lo = (int)(b & 0xFFFFFFFFll), hi = (int)(b >> 32);
    const double a0 = __longlong_as_double(((long long)__builtin_amdgcn_readlane(hi, 0) << 32) |
                                           (unsigned int)__builtin_amdgcn_readlane(lo, 0));
    const double a1 = __longlong_as_double(((long long)__builtin_amdgcn_readlane(hi, 32) << 32) |
                                           (unsigned int)__builtin_amdgcn_readlane(lo, 32));
    v = a0 + a1;
  }
  return v;
}

template <int G>
__device__ __forceinline__ DD grp_sum_dd(DD v) {
  v = dd_add(v, DD{xlane_d<1>(v.hi), xlane_d<1>(v.lo)});
  v = dd_add(v, DD{xlane_d<2>(v.hi), xlane_d<2>(v.lo)});
  if constexpr (G >= 8) v = dd_add(v, DD{xlane_d<4>(v.hi), xlane_d<4>(v.lo)});
  if constexpr (G >= 16) v = dd_add(v, DD{xlane_d<8>(v.hi), xlane_d<8>(v.lo)});
  if constexpr (G >= 32) v = dd_add(v, DD{xlane_d<16>(v.hi), xlane_d<16>(v.lo)});
  if constexpr (G >= 64) v = dd_add(v, DD{__shfl_xor(v.hi, 32), __shfl_xor(v.lo, 32)});
  return v;
}

// Per-chunk context of an owner row u.
template <bool W>
struct Chunk {
  Table<W> tab;
  QTable<W> qtab;
  int32_t u;
  int du;
  int32_t lo, hi;   // id range of the chunk's columns
  int c, nch;       // chunk index / count
  int64_t kbase;    // position of task entry j is kbase + j
  double su;        // row sum of u used by the epilogue
  // multi-chunk scratch, indexed by item: resume position and partial result
  int* cursor;
  uint32_t* pcnt;
  DD* pacc;
  // bitmap table (rows whose id span fits the LDS): bit (y - lo) is set for
  // every key y. Unweighted: word i at bits[i]. Weighted: word i at bits[2i]
  // and 1 + the number of keys in the words before it at bits[2i + 1] (one
  // 8-byte read per probe); bval[1 + r] = weight of the r-th key in id order
  // (as fp64, exact) and bval[0] = 0, the slot a missing lane reads. Word
  // wzero (one past the last) is zero: ids outside the span read it.
  const uint32_t* bits;
  const double* bval;
  uint32_t wzero;
};

// Probes U elements per lane against a bitmap table: one LDS read per probe
// (weighted: the word and its rank together), then weighted tables read the
// weight at rank + keys below the bit in the word, or the zero word. An id
// outside the row's span reads the zero guard word (ids below lo wrap to
// large offsets); bits past the span in the last word are zero, so no range
// test is needed.
template <bool W, int U, bool FAST>
__device__ __forceinline__ void probe_bm(const Chunk<W>& cx, const int32_t (&y)[U], int& cnt, DD& acc) {
  uint32_t off[U], wd[U], rk[U];
#pragma unroll
  for (int t = 0; t < U; ++t) {
    off[t] = (uint32_t)(y[t] - cx.lo);
    const uint32_t wi = min(off[t] >> 5, cx.wzero);
    if constexpr (W) {
      const uint2 p = reinterpret_cast<const uint2*>(cx.bits)[wi];
      wd[t] = p.x;
      rk[t] = p.y;
    } else {
      wd[t] = cx.bits[wi];
      rk[t] = 0u;
    }
  }
#pragma unroll
  for (int t = 0; t < U; ++t) {
    // bit (off & 31) of the word; v_bfe reads the low five bits of its
    // offset and width operands
    if constexpr (W) {
      // 0 or all ones
      const uint32_t hitm = (uint32_t)__builtin_amdgcn_sbfe((int)wd[t], off[t], 1u);
      cnt -= (int)hitm;
      // slot 1 + rank of the key (keys of the word below the bit counted
      // onto the word's stored 1 + rank), or slot 0 on a miss
      const uint32_t r = __builtin_amdgcn_ubfe(wd[t], 0u, off[t]);
      const double v = cx.bval[((uint32_t)__popc(r) + rk[t]) & hitm];
      if constexpr (FAST)
        acc.hi += v;
      else
        acc = dd_add1(acc, v);
    } else {
      cnt += (int)__builtin_amdgcn_ubfe(wd[t], off[t], 1u);
    }
  }
}

// Probes U elements per lane: per-lane hit count and weight sum. FAST: the
// weight sums are known to be exact in plain fp64 (see ExactInfo), so hits
// are added with one fp64 add instead of a double-double add; the result is
// the same.
//
// FA: all buckets of the pass are read before the first compare (one LDS
// round trip for the keys of a pass instead of one per probe, for more
// registers); the workgroup kernel uses it, the one-wave-per-row kernel,
// whose occupancy the extra registers would cut, does not.
template <bool W, int U, bool FAST, bool FA>
__device__ __forceinline__ void probe_u(const Chunk<W>& cx, const int32_t (&y)[U], int& cnt, DD& acc) {
  uint32_t bk[U];
  Hit h[U];
  bool anyneed = false;
  if constexpr (FA) {
    int4 ka[U], kb[U];
#pragma unroll
    for (int t = 0; t < U; ++t) {
      bk[t] = cx.tab.bucket(y[t]);
      cx.tab.fetch(bk[t], ka[t], kb[t]);
    }
#pragma unroll
    for (int t = 0; t < U; ++t) {
      h[t] = cx.tab.match(bk[t], y[t], ka[t], kb[t]);
      anyneed |= h[t].spill && !h[t].hit;
    }
  } else {
#pragma unroll
    for (int t = 0; t < U; ++t) {
      bk[t] = cx.tab.bucket(y[t]);
      h[t] = cx.tab.look(bk[t], y[t]);
      anyneed |= h[t].spill && !h[t].hit;
    }
  }
  if (ballot(anyneed)) {
#pragma unroll
    for (int t = 0; t < U; ++t) {
      while (h[t].spill && !h[t].hit) {
        bk[t] = cx.tab.next(bk[t]);
        h[t] = cx.tab.look(bk[t], y[t]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < U; ++t) {
    cnt += h[t].hit ? 1 : 0;
    if constexpr (W) {
      if constexpr (FAST)
        acc.hi += (double)h[t].v;
      else
        acc = dd_add1(acc, (double)h[t].v);
    }
  }
}

// probe_u for a quotient table: one 16-byte bucket read per element, all of
// the pass's buckets read before the first compare. A miss in a bucket whose
// spill flag is set looks at the next bucket with the displacement one
// higher, up to the table's largest displacement.
template <bool W, int U, bool FAST>
__device__ __forceinline__ void probe_q(const Chunk<W>& cx, const int32_t (&y)[U], int& cnt, DD& acc) {
  const QTable<W>& tb = cx.qtab;
  uint32_t bk[U], key[U];
  uint4 wv[U];
  Hit h[U];
  bool anyneed = false;
#pragma unroll
  for (int t = 0; t < U; ++t) {
    tb.split(y[t], bk[t], key[t]);
    wv[t] = tb.fetch(bk[t]);
  }
#pragma unroll
  for (int t = 0; t < U; ++t) {
    h[t] = tb.match(key[t], wv[t]);
    anyneed |= h[t].spill && !h[t].hit;
  }
  if (ballot(anyneed)) {
#pragma unroll
    for (int t = 0; t < U; ++t) {
      uint32_t k16 = key[t];
      const uint32_t step = 1u << tb.rb;
      const uint32_t last = key[t] + ((uint32_t)tb.maxd << tb.rb);
      while (h[t].spill && !h[t].hit && k16 < last) {
        bk[t] = (bk[t] + 1) & tb.bmask;
        k16 += step;
        h[t] = tb.match(k16, tb.fetch(bk[t]));
      }
    }
  }
#pragma unroll
  for (int t = 0; t < U; ++t) {
    cnt += h[t].hit ? 1 : 0;
    if constexpr (W) {
      if constexpr (FAST)
        acc.hi += (double)h[t].v;
      else
        acc = dd_add1(acc, (double)h[t].v);
    }
  }
}

// Probes the U entries of one pass per lane (bitmap or hash table). COUNT:
// also counts the entries below u (u's position in the row, the mirror
// position; whether u is there is checked when the pair is emitted): one
// compare per entry, counted per lane, or for whole-wave groups by the
// scalar unit from the compare's lane mask.
template <bool W, int G, int U, bool FAST, bool COUNT, bool STOP, int TK, bool FA>
__device__ __forceinline__ void probe_pass(const Chunk<W>& cx, int32_t (&y)[U], int& cnt, DD& acc, int& below) {
  if constexpr (COUNT) {
#pragma unroll
    for (int t = 0; t < U; ++t) {
      if constexpr (G == kWave)
        below += popc64(ballot(y[t] < cx.u));
      else
        below += y[t] < cx.u ? 1 : 0;
    }
  }
  if constexpr (TK == kTkBitmap)
    probe_bm<W, U, FAST>(cx, y, cnt, acc);
  else if constexpr (TK == kTkQuot)
    probe_q<W, U, FAST>(cx, y, cnt, acc);
  else
    probe_u<W, U, FAST, FA>(cx, y, cnt, acc);
}

// STOP: position of the first entry of the pass at p0 above cx.hi, or -1 when
// the pass does not cross the owner chunk's id bound (rows are sorted: it does
// iff its last entry does).
template <int G, int U>
__device__ __forceinline__ int pass_stop(const int32_t (&y)[U], int32_t hi, int p0, int e) {
  const uint64_t gm = group_mask<G>();
  if (!(ballot(y[U - 1] > hi) & gm)) return -1;
  const int gb = lane_id() & ~(G - 1);
  int first = e;
#pragma unroll
  for (int t = U - 1; t >= 0; --t) {
    const uint64_t bm = ballot(y[t] > hi) & gm;
    if (bm) first = p0 + t * G + (__ffsll((unsigned long long)bm) - 1 - gb);
  }
  return min(first, e);
}

// Word 3 of a raw buffer resource on gfx9-family parts: 32-bit data format,
// untyped dword loads.
constexpr int kBufWord3 = 0x00020000;

// Loads one pass: lane gl of the group reads entries p + t*G + gl, t < U.
// BUF: raw buffer loads through one resource over ind with 32-bit byte
// offsets (the launcher picks this form when 4 * m fits them). A read at or
// past the end of ind returns 0, so a pass that runs past the row's end needs
// no clamping (its extra values are replaced before use) and each load is one
// offset add with the t*G*4 part in the instruction. Otherwise 64-bit
// addresses with the index clamped to the row.
template <bool BUF, int G, int U>
__device__ __forceinline__ void load_pass(const EngineArgs& a, int32_t (&y)[U], int p, int e) {
  const int gl = lane_id() & (G - 1);
  if constexpr (BUF) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int32_t*>(a.ind), (short)0, (int)(uint32_t)(a.m * 4), kBufWord3);
    const uint32_t vo = (uint32_t)(p + gl) * 4u;
#pragma unroll
    for (int t = 0; t < U; ++t)
      y[t] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(vo + (uint32_t)(t * G * 4)), 0, 0);
  } else {
#pragma unroll
    for (int t = 0; t < U; ++t) y[t] = a.ind[min(p + t * G + gl, e - 1)];
  }
}

// Streams entries [s, e) of a short row (bounds uniform in the group of G
// lanes), U loads per lane per pass. While a further pass exists its loads
// are issued before the current pass is probed (global latency overlaps the
// LDS probes); the loop issues them on every iteration, so the wait for the
// current entries never waits for the ones just issued. The last pass
// (partial: lanes past the row's end probe a key that is never stored) runs
// after the loop. COUNT: also count the entries below u and whether u occurs
// (mirror position). STOP: the owner's chunk ends at id cx.hi; stop after the
// pass that crosses it and return the position of the first entry above
// cx.hi (entries of that pass beyond it are probed too, and miss). Returns
// the stop position, or e.
template <bool W, int G, int U, bool FAST, bool COUNT, bool STOP, int TK, bool ONEPASS, bool BUF, bool FA>
__device__ __forceinline__ int stream_t(const EngineArgs& a, const Chunk<W>& cx, int s, int e, int& cnt, DD& acc,
                                        int& below) {
  const int gl = lane_id() & (G - 1);
  int p0 = s;
  if (p0 >= e) return e;
  int32_t y[U];
  load_pass<BUF, G, U>(a, y, p0, e);
  // ONEPASS: the caller knows every row fits one pass (e - s <= G * U).
  // Otherwise the loop runs while the pass at p0 is full and another
  // follows; it is unrolled twice so that the two register sets take turns
  // (loads into one while the other is probed) without copies.
  // unrolled for bitmap and weighted probes; measured slower for
  // unweighted hash-table probes
  constexpr bool kPingPong = JC_PINGPONG && (TK == kTkBitmap || W);
  if constexpr (!ONEPASS && !kPingPong) {
    while (p0 + G * U < e) {
      int32_t yn[U];
      load_pass<BUF, G, U>(a, yn, p0 + G * U, e);
      probe_pass<W, G, U, FAST, COUNT, STOP, TK, FA>(cx, y, cnt, acc, below);
      if constexpr (STOP) {
        const int st = pass_stop<G, U>(y, cx.hi, p0, e);
        if (st >= 0) return st;
      }
      p0 += G * U;
#pragma unroll
      for (int t = 0; t < U; ++t) y[t] = yn[t];
    }
  }
  if constexpr (!ONEPASS && kPingPong) {
    int32_t yb[U];
    for (;;) {
      if (p0 + G * U >= e) break;
      load_pass<BUF, G, U>(a, yb, p0 + G * U, e);
      probe_pass<W, G, U, FAST, COUNT, STOP, TK, FA>(cx, y, cnt, acc, below);
      if constexpr (STOP) {
        const int st = pass_stop<G, U>(y, cx.hi, p0, e);
        if (st >= 0) return st;
      }
      p0 += G * U;
      if (p0 + G * U >= e) {
#pragma unroll
        for (int t = 0; t < U; ++t) y[t] = yb[t];
        break;
      }
      load_pass<BUF, G, U>(a, y, p0 + G * U, e);
      probe_pass<W, G, U, FAST, COUNT, STOP, TK, FA>(cx, yb, cnt, acc, below);
      if constexpr (STOP) {
        const int st = pass_stop<G, U>(yb, cx.hi, p0, e);
        if (st >= 0) return st;
      }
      p0 += G * U;
    }
  }
#pragma unroll
  for (int t = 0; t < U; ++t)
    if (p0 + t * G + gl >= e) y[t] = INT32_MAX;
  probe_pass<W, G, U, FAST, COUNT, STOP, TK, FA>(cx, y, cnt, acc, below);
  if constexpr (STOP) {
    const int st = pass_stop<G, U>(y, cx.hi, p0, e);
    if (st >= 0) return st;
  }
  return e;
}

template <bool W, int G, int U, bool FAST, bool COUNT, bool STOP, int TK, bool ONEPASS, bool FA>
__device__ __forceinline__ int stream(const EngineArgs& a, const Chunk<W>& cx, int s, int e, int& cnt, DD& acc,
                                      int& below) {
  if (!a.wide_index)
    return stream_t<W, G, U, FAST, COUNT, STOP, TK, ONEPASS, true, FA>(a, cx, s, e, cnt, acc, below);
  return stream_t<W, G, U, FAST, COUNT, STOP, TK, ONEPASS, false, FA>(a, cx, s, e, cnt, acc, below);
}

// A unit's result for pair slot `slot`, kept in LDS until the task's emits
// (emit_deferred): the count (unweighted) or the exact sum (weighted, its
// leading double), and the mirror position (-1: u not in row x). The slots
// are the multi-chunk partial and cursor arrays, which a unit reads before it
// writes its own slot.
template <bool W>
__device__ __forceinline__ void defer_result(const Chunk<W>& cx, int slot, uint32_t cnt, const DD& acc, int64_t q) {
  if constexpr (W)
    cx.pacc[slot].hi = acc.hi;
  else
    cx.pcnt[slot] = cnt;
  cx.cursor[slot] = (int)q;
}

// The emits of a task's (or row's) pairs, one thread per pair once every unit
// has stored its result: the division and the two stores run with all lanes
// busy instead of one lane per unit. Pair i is row x = ix[i] (degree
// idx_[i]) at entry ij[i] of the owner row u, whose entry 0 is at kbase.
template <int MODE, bool W>
__device__ __forceinline__ void emit_deferred(const EngineArgs& a, int32_t u, int du, double su, int64_t kbase,
                                              int i, int32_t x, int dx, int j, const int* rq,
                                              const uint32_t* rcnt, const DD* racc) {
  const double sx = row_sum_of<MODE>(a, x);
  DD acc{0.0, 0.0};
  uint32_t cnt = 0;
  if constexpr (W)
    acc.hi = racc[i].hi;
  else
    cnt = rcnt[i];
  // the mirror position: entry `below` of row x, if that entry is u (a CSR
  // that is not symmetric may lack (x, u); that position is left to the
  // fix-up pass)
  int64_t q = rq[i];
  if (q >= 0 && a.ind[q] != u) q = -1;
  emit<MODE, W>(a, u, du, x, dx, kbase + j, q, cnt, acc, su, sx);
}

// One unit of work: each group of G lanes handles one owned short row x
// (valid), U loads per lane in flight.
//   MULTI = false (the owner row fits one chunk): stream all of N(x), probe,
//     count the entries below u (mirror position) and emit.
//   MULTI = true: from the row's resume position, stream entries up to the
//     chunk's id bound, store the partial and the resume position; on the
//     last chunk stream to the end, add the stored partial, locate u by one
//     group search and emit.
template <int MODE, bool W, int G, int U, bool MULTI, bool FAST, int TK, bool SMALL = false>
__device__ __forceinline__ void unit(const EngineArgs& a, const Chunk<W>& cx, bool valid, int32_t x,
                                     int ox, int dx, int j, int item) {
  const int gl = lane_id() & (G - 1);
  const uint64_t gm = group_mask<G>();
  const int end = ox + dx;
  const bool last = !MULTI || cx.c == cx.nch - 1;
  int s = ox;
  if constexpr (MULTI) s = (valid && cx.c > 0) ? cx.cursor[item] : ox;
  int e = valid ? end : s;
  if constexpr (G == kWave) {
    // one row per wave: the bounds are wave-uniform, so the streaming loop's
    // control runs on the scalar unit
    s = __builtin_amdgcn_readfirstlane(s);
    e = __builtin_amdgcn_readfirstlane(e);
  }
  // single-chunk units count the entries below u while streaming (the mirror
  // position); multi-chunk units locate u with one group search at the end
  constexpr bool kCountMirror = !MULTI && (JC_MIRRORCOUNT || G * U <= 64);
  constexpr bool kOnePass = G * U <= 64;
  constexpr bool kFetchAll = JC_FETCHALL && !SMALL;  // probe_u
  int cnt = 0, below = 0;
  DD acc{0.0, 0.0};
  int stop_at = e;
  if (MULTI && !last)
    stop_at = stream<W, G, U, FAST, kCountMirror, true, TK, false, kFetchAll>(a, cx, s, e, cnt, acc, below);
  else
    stream<W, G, U, FAST, kCountMirror, false, TK, SMALL || (!MULTI && kOnePass), kFetchAll>(a, cx, s, e, cnt, acc,
                                                                                            below);
  (void)gm;
  (void)gl;
  if constexpr (kCountMirror && kOnePass) {
    // one pass of at most U entries per lane: both counts fit a byte each
    const int packed = grp_sum_i<G>(cnt | (below << 8));
    cnt = packed & 0xFF;
    below = packed >> 8;
  } else if constexpr (kCountMirror) {
    cnt = grp_sum_i<G>(cnt);
    if constexpr (G != kWave) below = grp_sum_i<G>(below);  // whole-wave counts are totals already
  } else {
    cnt = grp_sum_i<G>(cnt);
  }
  if constexpr (W) {
    if constexpr (FAST)
      acc.hi = grp_sum_f64<G>(acc.hi);
    else
      acc = grp_sum_dd<G>(acc);
  }
  if constexpr (!MULTI) {
    int64_t q = -1;
    if constexpr (kCountMirror) {
      // u sits at entry `below` of row x if it is in the row (the emit checks)
      if (below < dx) q = (int64_t)ox + below;
    } else {
      int pos = end;
      if (valid) pos = grp_lower_bound<G>(a.ind, ox, end, cx.u);
      if (valid && pos < end && a.ind[pos] == cx.u) q = pos;
    }
    if (valid && gl == 0) defer_result<W>(cx, SMALL ? j : item, (uint32_t)cnt, acc, q);
  } else {
    if (!last) {
      if (valid && gl == 0) {
        cx.cursor[item] = stop_at;
        if constexpr (W) {
          // FAST: partial sums are exact in plain fp64 too
          if constexpr (FAST)
            cx.pacc[item] = DD{cx.c == 0 ? acc.hi : cx.pacc[item].hi + acc.hi, 0.0};
          else
            cx.pacc[item] = cx.c == 0 ? acc : dd_add(cx.pacc[item], acc);
        } else {
          cx.pcnt[item] = cx.c == 0 ? (uint32_t)cnt : cx.pcnt[item] + (uint32_t)cnt;
        }
      }
    } else {
      int q = end;
      if (valid) q = grp_lower_bound<G>(a.ind, ox, end, cx.u);
      if (valid && gl == 0) {
        uint32_t tot = (uint32_t)cnt;
        if constexpr (W) {
          if constexpr (FAST)
            acc.hi += cx.pacc[item].hi;
          else
            acc = dd_add(cx.pacc[item], acc);
        } else {
          tot += cx.pcnt[item];
        }
        const bool f = q < end && a.ind[q] == cx.u;
        defer_result<W>(cx, item, tot, acc, f ? (int64_t)q : -1);
      }
    }
  }
}

// Length classes of a short row (entries streamed per chunk):
//   0: <= 8     G 8,  U 1      3: <= 64   G 16, U 4
//   1: <= 16    G 8,  U 2      4: <= 512  G 32, U 4
//   2: <= 32    G 8,  U 4      5: > 512   G 64, U 4 (weighted bitmap
//                                          rows: 3; unweighted bitmap rows: 6)
// (classes 1 and 2 in the one-wave-per-row kernel: G 16, U 1 and 2)
// (Class 4 up to 512 entries rather than 256: two rows per wave in flight
// instead of one, measured 58.7 -> 57.2 ms on RMAT-22 weighted, 3.45 -> 3.35
// ms on Chung-Lu 2^20, 194 -> 187 ms for the batched step; 384 to 768 were
// within 0.1% of each other, 1024 and above slower on RMAT-22.)
#ifndef JC_U4
#define JC_U4 4
#endif
#ifndef JC_U5
#define JC_U5 4
#endif
#ifndef JC_U5W
#define JC_U5W 4  // four loads per lane: best once rows of up to 512 entries went to class 4 (3 before)
#endif
// the same for rows probed against a bitmap table; the cheaper unweighted
// bitmap probe keeps more loads in flight (six measured best on the batched
// workload: 2 and 3 slower, 8 no better)
#ifndef JC_U4BM
#define JC_U4BM JC_U4
#endif
#ifndef JC_U5BM
#define JC_U5BM 6
#endif
#ifndef JC_U5WBM
#define JC_U5WBM 3  // weighted bitmap rows: three loads per lane
#endif
#ifndef JC_C4_MAX
#define JC_C4_MAX 512
#endif
constexpr int kClasses = 6;
__host__ __device__ __forceinline__ int len_class(int len) {
  return len <= 8 ? 0 : len <= 16 ? 1 : len <= 32 ? 2 : len <= 64 ? 3 : len <= JC_C4_MAX ? 4 : 5;
}
// Lanes per group of a class. The workgroup kernel streams rows of up to 32
// entries with 8 lanes (8 rows per wave; 16 lanes measured 0.8% slower on
// RMAT-22 weighted and 3% on Chung-Lu 2^20); the one-wave-per-row kernel
// (small) keeps 16 lanes for classes 1-2, where 8 measured 4% slower.
__host__ __device__ __forceinline__ constexpr int class_g(int c, bool small = false) {
  return (c == 0 || (!small && c <= 2)) ? 8 : c <= 3 ? 16 : c == 4 ? 32 : 64;
}

template <int MODE, bool W, bool MULTI, int C, int TK = kTkHash, bool SMALL = false>
__device__ __forceinline__ void unit_class(const EngineArgs& a, const Chunk<W>& cx, bool valid, int32_t x,
                                           int ox, int dx, int j, int item, bool fast) {
  constexpr int G = class_g(C, SMALL);
  constexpr int U = C == 0 ? 1
                  : C == 1 ? (G == 8 ? 2 : 1)
                  : C == 2 ? (G == 8 ? 4 : 2)
                  : C == 3 ? 4
                  : C == 4 ? (TK == kTkBitmap ? JC_U4BM : JC_U4)
                           : (W ? (TK == kTkBitmap ? JC_U5WBM : JC_U5W) : (TK == kTkBitmap ? JC_U5BM : JC_U5));
  // classes 0-3 stream a row (at most 64 entries) in one pass of G * U; the
  // single-pass mirror count and its byte-packed reduction rely on it
  static_assert(C > 3 ? G * U > 64 : G * U >= (C == 0 ? 8 : C == 1 ? 16 : C == 2 ? 32 : 64), "class shape");
  // rows of the one-wave-per-row kernel (<= kSmallMax entries) fit one pass
  static_assert(!SMALL || C > 4 || G * U >= (C == 4 ? kSmallMax : 0), "small rows fit one pass");
#if JC_F64FAST
  if constexpr (W) {
    if (fast) {
      unit<MODE, W, G, U, MULTI, true, TK, SMALL>(a, cx, valid, x, ox, dx, j, item);
      return;
    }
  }
#endif
  (void)fast;
  unit<MODE, W, G, U, MULTI, false, TK, SMALL>(a, cx, valid, x, ox, dx, j, item);
}

// Items of one class held in lanes (x, ox, dx, j per lane; mask = owned lanes
// of the class): processed 64/G at a time, group g taking the g-th item.
template <int MODE, bool W, int C>
__device__ __forceinline__ void run_class_regs(const EngineArgs& a, const Chunk<W>& cx, uint64_t mask,
                                               int32_t x, int ox, int dx, int j, bool fast) {
  constexpr int G = class_g(C, true);
  constexpr int NG = kWave / G;
  const int g = lane_id() / G;
  while (mask) {
    int src = -1;
#pragma unroll
    for (int i = 0; i < NG; ++i) {
      if (mask) {
        const int b = __ffsll((unsigned long long)mask) - 1;
        if (i == g) src = b;
        mask &= mask - 1;
      }
    }
    const int s = src < 0 ? 0 : src;
    const int32_t xx = __shfl(x, s);
    const int oxx = __shfl(ox, s);
    const int dxx = __shfl(dx, s);
    const int jj = __shfl(j, s);
    // rows here have at most kSmallMax entries: one pass of every class of
    // such rows covers them (class 4: G * U = 128)
    unit_class<MODE, W, false, C, kTkHash, true>(a, cx, src >= 0, xx, oxx, dxx, jj, 0, fast);
  }
}

// Candidate neighbours in lanes: split the owned ones by short-row length and
// run each class (single-chunk owner rows).
template <int MODE, bool W>
__device__ __forceinline__ void run_batch_regs(const EngineArgs& a, const Chunk<W>& cx, bool cand, int32_t x,
                                               int ox, int dx, int j, bool fast) {
  const bool own = cand && owns(cx.u, cx.du, x, dx);
  const int c = len_class(dx);
  uint64_t m[kClasses];
#pragma unroll
  for (int k = 0; k < kClasses; ++k) m[k] = ballot(own && c == k);
  if (m[5]) run_class_regs<MODE, W, 5>(a, cx, m[5], x, ox, dx, j, fast);
  if (m[4]) run_class_regs<MODE, W, 4>(a, cx, m[4], x, ox, dx, j, fast);
  if (m[3]) run_class_regs<MODE, W, 3>(a, cx, m[3], x, ox, dx, j, fast);
  if (m[2]) run_class_regs<MODE, W, 2>(a, cx, m[2], x, ox, dx, j, fast);
  if (m[1]) run_class_regs<MODE, W, 1>(a, cx, m[1], x, ox, dx, j, fast);
  if (m[0]) run_class_regs<MODE, W, 0>(a, cx, m[0], x, ox, dx, j, fast);
}

// ---------------------------------------------------------------------------
// Rows with degree <= kSmallMax: one wave per row, a private table per wave.
template <int MODE, bool W, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_owner_small(EngineArgs a) {
  constexpr int NW = BLOCK / kWave;
  constexpr int SLOTS = kBucket << 6;  // 64 buckets hold a row of up to 128 keys at <= 2 per bucket
  __shared__ __align__(16) int32_t key_all[NW][SLOTS];
  __shared__ __align__(16) float val_all[W ? NW : 1][W ? SLOTS : 1];
  // per-wave results of a row's pairs until the row's emits, by entry index
  __shared__ int s_rq[NW][kSmallMax];
  __shared__ __align__(16) uint32_t s_rcnt[W ? 1 : NW][W ? 1 : kSmallMax];
  __shared__ __align__(16) DD s_racc[W ? NW : 1][W ? kSmallMax : 1];
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int nsmall = *a.num_small;
  const bool fast = exact_f64(a);
  const int stride = gridDim.x * NW;
  for (int r = blockIdx.x * NW + w; r < nsmall; r += stride) {
    const int32_t u = a.small_rows[r];
    const int ou = a.off[u];
    const int du = a.off[u + 1] - ou;
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
    const bool c0 = lane < du, c1 = lane + kWave < du;
    if ((ballot(c0 && owns(u, du, x0, dx0)) | ballot(c1 && owns(u, du, x1, dx1))) == 0) continue;
    const int lb = table_log_buckets<W>(du);
    Chunk<W> cx;
    cx.tab = Table<W>{key_all[w], W ? val_all[W ? w : 0] : nullptr, lb, (uint32_t)((1 << lb) - 1), SLOTS / 2};
    for (int i = lane; i < (4 << lb); i += kWave) cx.tab.clear(i);
    if (lane == 0) cx.tab.set_zero();
    lds_fence();
    build_table<W>(cx.tab, a.ind, a.vw, ou, du, lane, kWave);
    lds_fence();
    cx.u = u;
    cx.du = du;
    cx.su = row_sum_of<MODE>(a, u);
    cx.lo = a.ind[ou];
    cx.hi = a.ind[ou + du - 1];
    cx.c = 0;
    cx.nch = 1;
    cx.kbase = ou;
    cx.cursor = s_rq[w];
    cx.pcnt = W ? nullptr : s_rcnt[W ? 0 : w];
    cx.pacc = W ? s_racc[W ? w : 0] : nullptr;
    run_batch_regs<MODE, W>(a, cx, c0, x0, ox0, dx0, lane, fast);
    run_batch_regs<MODE, W>(a, cx, c1, x1, ox1, dx1, lane + kWave, fast);
    lds_fence();  // the units' results are visible to every lane
    // each lane emits the pairs of its own entries (j = lane, lane + 64)
    if (c0 && owns(u, du, x0, dx0))
      emit_deferred<MODE, W>(a, u, du, cx.su, ou, lane, x0, dx0, lane, cx.cursor, cx.pcnt, cx.pacc);
    if (c1 && owns(u, du, x1, dx1))
      emit_deferred<MODE, W>(a, u, du, cx.su, ou, lane + kWave, x1, dx1, lane + kWave, cx.cursor, cx.pcnt, cx.pacc);
    lds_fence();  // all probes and emits of this row done before the next row clears the table
  }
}

// ---------------------------------------------------------------------------
// Rows with degree > kSmallMax: one workgroup per (row, piece) task.
// LDS of one task workgroup (one per CU, the whole 160 KiB): what the item
// list and per-item state leave is the hash table. Measured on RMAT-22 this
// beats two 80 KiB workgroups per CU: halving the chunk count of hub rows
// outweighs the idle time at chunk barriers.
#ifndef JC_BLOCK_LDS_KIB
#define JC_BLOCK_LDS_KIB 160
#endif
#ifndef JC_BLOCK
#define JC_BLOCK 1024
#endif
constexpr size_t kBlockLds = JC_BLOCK_LDS_KIB * 1024 - 1024;  // minus the kernel's static shared variables

#ifndef JC_BITMAP
#define JC_BITMAP 1
#endif

// Table kind of a long row, from its first and last column: the bitmap when
// its words and the guard word (weighted: with their ranks) and, weighted,
// the zero slot and the keys' fp64 weights fit the table region of the LDS,
// and, weighted, the words and ranks take at most kBitmapWordsMaxW bytes:
// a weighted bitmap over a wider id span (a sparse row) is cleared, ranked
// and probed more slowly than a quotient table of its keys (measured: RMAT
// and Chung-Lu graphs of 2^18 vertices 26-30% faster on quotient tables,
// 2^14-2^16 vertices 2-2.5x faster on bitmaps, 2^17 even; unweighted rows
// are faster on bitmaps at every span that fits). Used by the scheduler and
// by the task kernels alike.
constexpr size_t kBitmapWordsMaxW = 32 * 1024;
template <bool W>
__host__ __device__ __forceinline__ bool bitmap_fits(int32_t lo, int32_t hi, int du, size_t budget) {
  const uint32_t words = ((uint32_t)(hi - lo) >> 5) + 1u;  // plus one zero guard word
  const size_t word_bytes = (size_t)(words + 1u) * (W ? 8u : 4u);
  return JC_BITMAP && word_bytes + (W ? (size_t)8 * (du + 1) : 0) <= budget &&
         (!W || word_bytes <= kBitmapWordsMaxW);
}

// Exclusive prefix sum over the workgroup (all threads active).
__device__ __forceinline__ int block_excl_scan(int v, int* s_wsum) {
  const int lane = lane_id(), w = threadIdx.x / kWave;
  int incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == kWave - 1) s_wsum[w] = incl;
  __syncthreads();
  int base = 0;
  for (int i = 0; i < w; ++i) base += s_wsum[i];
  __syncthreads();
  return base + incl - v;
}

// Table region of the bitmap-task kernel when two workgroups share a CU:
// 2 x (32 KiB + item list + staging + results + static LDS) fits the 160 KiB
// of a CU.
constexpr size_t kBitmapLds2 = 32 * 1024;

template <bool W>
struct BlockCfg {
  static constexpr size_t item_bytes = (size_t)kTaskSpan * (3 * sizeof(int32_t) + sizeof(uint16_t));
  static constexpr size_t cursor_bytes = (size_t)kTaskSpan * sizeof(int);
  static constexpr size_t part_bytes = (size_t)kTaskSpan * (W ? sizeof(DD) : sizeof(uint32_t));
  static constexpr size_t bucket_bytes = kBucket * (sizeof(int32_t) + (W ? sizeof(float) : 0));
  static constexpr int NB_FIT = (int)((kBlockLds - item_bytes - cursor_bytes - part_bytes) / bucket_bytes);
  static constexpr int NB = 1 << (31 - __builtin_clz(NB_FIT));  // power of two
  static constexpr int CH = kLambda<W> * NB;  // keys per chunk
  static constexpr size_t key_bytes = (size_t)NB * kBucket * sizeof(int32_t);
  static constexpr size_t val_bytes = W ? key_bytes : 0;
  static constexpr size_t lds_bytes = key_bytes + val_bytes + item_bytes + cursor_bytes + part_bytes;
  // bitmap tasks: staging of the next task's entries (column, its offset and
  // degree, and weighted, its vertex weight)
  static constexpr size_t stage_bytes = (size_t)kTaskSpan * (3 * sizeof(int32_t) + (W ? sizeof(float) : 0));
  // bitmap tasks: everything but the table region, and that region with one
  // workgroup per CU
  static constexpr size_t bm_rest = item_bytes + stage_bytes + cursor_bytes + part_bytes;
  static constexpr size_t bm_budget1 = kBlockLds - bm_rest;
  // quotient tables: 16 bytes per bucket, weighted also kQLambda dense
  // weights per bucket and the zero slot
  // quotient-table tasks also stage the next piece's entries (column,
  // offset, degree), which leaves the bucket count unchanged
  static constexpr size_t qstage_bytes = (size_t)kTaskSpan * 3 * sizeof(int32_t);
  static constexpr size_t q_avail = kBlockLds - item_bytes - qstage_bytes - cursor_bytes - part_bytes;
  static constexpr size_t q_bucket_cost = 16 + (W ? 4 * kQLambda : 0);
  static constexpr int QNB_FIT = (int)((q_avail - 16) / q_bucket_cost);
  static constexpr int QNB = 1 << (31 - __builtin_clz(QNB_FIT));  // power of two
  static constexpr int QCH = kQLambda * QNB;                      // keys per chunk
  static constexpr size_t qword_bytes = (size_t)QNB * 16;
  static constexpr size_t qval_bytes = W ? ((size_t)(QCH + 1) * sizeof(float) + 15) / 16 * 16 : 0;
  static constexpr size_t q_lds_bytes =
      qword_bytes + qval_bytes + item_bytes + qstage_bytes + cursor_bytes + part_bytes;
  static_assert(QNB <= (1 << 14) && QCH < (1 << 15), "quotient table: lb <= 14, weight bases fit 15 bits");
  static_assert(QNB >= (1 << kQMinLb), "quotient table smaller than its minimum");
};

// Slots 0..6 of a quotient-table bucket that hold a key.
__device__ __forceinline__ int qbucket_count(const uint4& w) {
  return ((w.x & 0xFFFFu) != 0xFFFFu) + ((w.x >> 16) != 0xFFFFu) + ((w.y & 0xFFFFu) != 0xFFFFu) +
         ((w.y >> 16) != 0xFFFFu) + ((w.z & 0xFFFFu) != 0xFFFFu) + ((w.z >> 16) != 0xFFFFu) +
         ((w.w & 0xFFFFu) != 0xFFFFu);
}

// Builds a quotient table from keys ind[cs .. cs+len) (len <= KPT * BLOCK),
// weighted with vw[key]: inserts (four keys per thread loaded before any
// insert), then for weighted tables the buckets' weight bases (one prefix
// sum over the buckets) and the weights. A key that cannot be placed sets
// *s_fail (which the caller cleared before the build); every thread must
// call this.
template <bool W, int BLOCK, int KPT>
__device__ __forceinline__ void build_qtable(const QTable<W>& tab, const int32_t* __restrict__ ind,
                                             const float* __restrict__ vw, int cs, int len, int tid, int* s_fail,
                                             int* s_wsum) {
  static_assert(KPT % 4 == 0, "keys per thread in groups of four");
  constexpr uint32_t kNone = 0xFFFFFFFFu;
  uint32_t pos[W ? KPT : 1];
  float v[W ? KPT : 1];
  if constexpr (W) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) pos[k] = kNone;
  }
  bool ok = true;
#pragma unroll
  for (int k0 = 0; k0 < KPT; k0 += 4) {
    if (k0 * BLOCK >= len) break;  // uniform over the workgroup
    int32_t y[4];
    float vv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; ++t) y[t] = ind[cs + min(tid + (k0 + t) * BLOCK, len - 1)];
    if constexpr (W) {
#pragma unroll
      for (int t = 0; t < 4; ++t) vv[t] = vw[y[t]];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      uint32_t p = kNone;
      if (tid + (k0 + t) * BLOCK < len) {
        if (!tab.insert(y[t], p)) {
          ok = false;
          p = kNone;
        }
      }
      if constexpr (W) {
        pos[k0 + t] = p;
        v[k0 + t] = vv[t];
      }
    }
  }
  if (!ok) atomicOr(s_fail, 1);
  if constexpr (W) {
    // key counts of this thread's buckets, the block prefix sum, the bases
    __syncthreads();
    const int nb = 1 << tab.lb;
    const int per = (nb + BLOCK - 1) / BLOCK;
    const int b0 = min(nb, tid * per), b1 = min(nb, b0 + per);
    int c = 0;
    for (int b = b0; b < b1; ++b) c += qbucket_count(tab.fetch(b));
    int r = block_excl_scan(c, s_wsum);
    for (int b = b0; b < b1; ++b) {
      const int cb = qbucket_count(tab.fetch(b));
      tab.word[4 * b + 3] |= (uint32_t)r << 16;
      r += cb;
    }
    if (tid == 0) tab.val[0] = 0.0f;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      if (pos[k] != kNone) {
        const uint32_t b = pos[k] >> 3, sl = pos[k] & 7u;
        tab.val[1 + __builtin_amdgcn_ubfe(tab.word[4 * b + 3], 16, 15) + sl] = v[k];
      }
    }
  }
}

template <int MODE, bool W, bool MULTI, int TK = kTkHash>
__device__ __forceinline__ void run_unit_lds(const EngineArgs& a, const Chunk<W>& cx, int cls, int first,
                                             int cend, const int32_t* ix, const int32_t* iox,
                                             const int32_t* idx_, const uint16_t* ij, bool fast) {
  const int G = class_g(cls);
  const int item = first + lane_id() / G;
  const bool valid = item < cend;
  const int it = valid ? item : first;
  const int32_t x = ix[it];
  const int ox = iox[it];
  const int dx = idx_[it];
  const int j = ij[it];
  switch (cls) {
    case 5: unit_class<MODE, W, MULTI, 5, TK>(a, cx, valid, x, ox, dx, j, it, fast); break;
    case 4: unit_class<MODE, W, MULTI, 4, TK>(a, cx, valid, x, ox, dx, j, it, fast); break;
    case 3: unit_class<MODE, W, MULTI, 3, TK>(a, cx, valid, x, ox, dx, j, it, fast); break;
    case 2: unit_class<MODE, W, MULTI, 2, TK>(a, cx, valid, x, ox, dx, j, it, fast); break;
    case 1: unit_class<MODE, W, MULTI, 1, TK>(a, cx, valid, x, ox, dx, j, it, fast); break;
    default: unit_class<MODE, W, MULTI, 0, TK>(a, cx, valid, x, ox, dx, j, it, fast); break;
  }
}

// KIND: the task list and table this instantiation serves: 0 hash-table
// tasks with the 32-bit table, 1 bitmap tasks, 2 hash-table tasks with the
// quotient table (graphs below 2^24 vertices). A kind-2 task whose table
// cannot be built (a key finds no slot within the largest displacement) is
// abandoned before it writes anything and listed in a.redo; a kind-0 launch
// with a.redo_pass set then runs exactly those tasks with the 32-bit table.
// Separate instantiations keep the register allocation of one path from
// constraining the others.
template <int MODE, bool W, int BLOCK, int KIND>
__global__ __launch_bounds__(BLOCK) void k_owner_block(EngineArgs a) {
  static_assert(BLOCK >= kTaskSpan, "one thread per task entry during compaction");
  using Cfg = BlockCfg<W>;
  constexpr bool kQ = KIND == 2;
  constexpr int CH = kQ ? Cfg::QCH : Cfg::CH;
  constexpr int KPT = ((Cfg::QCH + BLOCK - 1) / BLOCK + 3) / 4 * 4;  // quotient build: keys per thread
  constexpr int NW = BLOCK / kWave;
  // LDS: the table region, then the item list; hash-table tasks add the
  // multi-chunk cursors and partials after it. Bitmap tasks are single-chunk,
  // and their table region is a.bm_budget bytes (the launcher sizes the
  // dynamic LDS of this kind to it plus the item list, so a smaller budget
  // lets two workgroups share a CU).
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* tkey = reinterpret_cast<int32_t*>(smem);
  float* tval = reinterpret_cast<float*>(smem + Cfg::key_bytes);
  uint32_t* qword = reinterpret_cast<uint32_t*>(smem);
  float* qval = reinterpret_cast<float*>(smem + Cfg::qword_bytes);
  unsigned char* items = smem + (KIND == 1 ? (size_t)a.bm_budget
                                 : kQ      ? Cfg::qword_bytes + Cfg::qval_bytes
                                           : Cfg::key_bytes + Cfg::val_bytes);
  int32_t* ix = reinterpret_cast<int32_t*>(items);
  int32_t* iox = ix + kTaskSpan;
  int32_t* idx_ = iox + kTaskSpan;
  uint16_t* ij = reinterpret_cast<uint16_t*>(idx_ + kTaskSpan);
  // multi-chunk cursors and partials (hash-table tasks), which also hold the
  // units' results until the task's emits (all tasks); bitmap tasks keep them
  // after their staging area
  unsigned char* resp =
      items + Cfg::item_bytes + (KIND == 1 ? Cfg::stage_bytes : kQ ? Cfg::qstage_bytes : 0);
  int* cursor = reinterpret_cast<int*>(resp);
  unsigned char* partp = resp + Cfg::cursor_bytes;
  uint32_t* pcnt = reinterpret_cast<uint32_t*>(partp);
  DD* pacc = reinterpret_cast<DD*>(partp);
  // Bitmap and quotient-table tasks look ahead: while the other waves run a
  // piece's units, wave 0 stages the entries of the next piece (column,
  // offset, degree; weighted whole-row bitmap tasks also the columns' vertex
  // weights) in LDS after the item list, and at a task's last piece reserves
  // the next task and stages its header too, so a piece starts without a
  // global round trip.
  constexpr bool kLook = KIND != 0;
  int32_t* st_x = kLook ? reinterpret_cast<int32_t*>(items + Cfg::item_bytes) : nullptr;
  int32_t* st_ox = kLook ? st_x + kTaskSpan : nullptr;
  int32_t* st_dx = kLook ? st_x + 2 * kTaskSpan : nullptr;
  float* st_w = KIND == 1 && W ? reinterpret_cast<float*>(st_x + 3 * kTaskSpan) : nullptr;
  __shared__ int s_task;
  __shared__ int s_pre;      // the next piece's entries (at a task's start: and header) are staged
  __shared__ int s_rsv;      // quotient tasks: s_task was reserved by the lookahead
  __shared__ int s_hdr[7];   // staged header: u, ou, du, ka, first and last column of u, pieces
  __shared__ int s_next;
  __shared__ int s_wcnt[kClasses][NW];
  __shared__ int s_cbase[kClasses + 1];   // item ranges per class, heaviest first
  __shared__ int s_ubase[kClasses + 1];   // unit ranges per class
  __shared__ int s_wsum[NW];
  __shared__ int s_fail;     // quotient tables: a key found no slot
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int w = tid / kWave;
  const bool redo = KIND == 0 && a.redo_pass != 0;
  const int ntasks = KIND == 1 ? a.num_tasks[1] : redo ? *a.num_redo : a.num_tasks[0];
  const int tbase = KIND == 1 ? a.num_tasks[0] : 0;
  int* const counter = &a.task_counter[KIND == 1 ? 1 : redo ? 2 : 0];
  const bool fast = exact_f64(a);
  if (kLook && tid == 0) {
    s_task = atomicAdd(counter, 1);
    s_pre = 0;
    s_rsv = 1;
  }
  for (;;) {
    __syncthreads();
    if constexpr (KIND != 1) {
      if (tid == 0) {
        // a quotient task that failed before its lookahead ran reserved nothing
        if (!kQ || !s_rsv) {
          s_task = atomicAdd(counter, 1);
          if (kQ) s_pre = 0;
        }
        if (kQ) s_rsv = 0;
      }
      __syncthreads();
    }
    const int t = s_task;
    if (t >= ntasks) return;
    const bool pre = kLook && s_pre != 0;
    int32_t u, lo0, hi0;
    int ou, du, ka;
    int npieces = 1;  // pieces of kTaskSpan entries this task covers
    if (pre) {
      u = s_hdr[0];
      ou = s_hdr[1];
      du = s_hdr[2];
      ka = s_hdr[3];
      lo0 = s_hdr[4];
      hi0 = s_hdr[5];
      npieces = kQ ? s_hdr[6] : 1;  // bitmap tasks are one piece
    } else {
      const int2 tk = a.tasks[tbase + (redo ? a.redo[t] : t)];
      u = tk.x;
      ou = a.off[u];
      du = a.off[u + 1] - ou;
      // (u, -1): the whole row (hash-table rows of one chunk); (u, p): piece p
      ka = ou + max(tk.y, 0) * kTaskSpan;
      if (KIND != 1 && tk.y < 0) npieces = (du + kTaskSpan - 1) / kTaskSpan;
      // the row's first and last column: bitmap tasks only
      lo0 = KIND == 1 ? a.ind[ou] : 0;
      hi0 = KIND == 1 ? a.ind[ou + du - 1] : 0;
    }
    // bitmap tasks: the row's id span fits the table region of the LDS
    // (bitmap_fits, decided by the scheduler)
    const uint32_t words = ((uint32_t)(hi0 - lo0) >> 5) + 1u;
    // rows longer than one chunk: nch chunks of equal size (a full chunk
    // and a short last one would fill the first chunk's buckets to the
    // limit, so more of its keys overflow into neighbouring buckets and
    // more probes walk the overflow)
    const int nch = KIND == 1 ? 1 : (du + CH - 1) / CH;
    const int chk = (du + nch - 1) / nch;
    bool built = true;  // hash-table tasks: false once handed to the 32-bit redo pass
    for (int pc = 0; pc < npieces && built; ++pc) {
      if (pc > 0) {
        ka += kTaskSpan;
        __syncthreads();  // the previous piece's emits have read the item list
      }
      const int span = min(kTaskSpan, ou + du - ka);
      const bool pre_p = pc == 0 ? pre : (kQ && s_pre != 0);

      // owned entries of the piece, sorted into length classes (heaviest first)
      int32_t x = 0;
      int ox = 0, dx = 0, cls = -1;
      if (tid < span) {
        if (pre_p) {
          x = st_x[tid];
          ox = st_ox[tid];
          dx = st_dx[tid];
        } else {
          x = a.ind[ka + tid];
          ox = a.off[x];
          dx = a.off[x + 1] - ox;
        }
        if (owns(u, du, x, dx)) cls = len_class((dx + nch - 1) / nch);
      }
      int rank = 0;
#pragma unroll
      for (int k = 0; k < kClasses; ++k) {
        const uint64_t bm = ballot(cls == k);
        if (cls == k) rank = popc64(bm & ((1ull << lane) - 1));
        if (lane == 0) s_wcnt[k][w] = popc64(bm);
      }
      __syncthreads();
      if (tid == 0) {
        if (kQ) s_pre = 0;  // every thread has read the staged piece
        int ib = 0, ub = 0;
        for (int k = kClasses - 1; k >= 0; --k) {
          int c = 0;
          for (int v = 0; v < NW; ++v) c += s_wcnt[k][v];
          s_cbase[kClasses - 1 - k] = ib;
          s_ubase[kClasses - 1 - k] = ub;
          ib += c;
          ub += (c + (kWave / class_g(k)) - 1) / (kWave / class_g(k));
        }
        s_cbase[kClasses] = ib;
        s_ubase[kClasses] = ub;
      }
      __syncthreads();
      if (cls >= 0) {
        int at = s_cbase[kClasses - 1 - cls] + rank;
        for (int v = 0; v < w; ++v) at += s_wcnt[cls][v];
        ix[at] = x;
        iox[at] = ox;
        idx_[at] = dx;
        ij[at] = (uint16_t)tid;
      }
      const int nunits = s_ubase[kClasses];
      if constexpr (KIND == 1) {
        // the words and the zero guard word (weighted: interleaved with
        // 1 + their ranks), then, weighted, the zero slot and the keys'
        // weights in id order as fp64
        constexpr uint32_t WS = W ? 2u : 1u;  // stride of the words
        uint32_t* bits = reinterpret_cast<uint32_t*>(smem);
        double* bval = reinterpret_cast<double*>(bits + WS * (words + 1u));
        // words plus the zero guard word (weighted: with its rank slot)
        for (uint32_t i = tid; i < WS * (words + 1u); i += BLOCK) bits[i] = 0u;
        if (tid == 0) s_next = 0;
        __syncthreads();
        if (du <= kTaskSpan) {
          // the task is the whole row: its entries are the keys, in id order
          if (tid < du) {
            const uint32_t o = (uint32_t)(x - lo0);
            if (o < words * 32u) atomicOr(&bits[WS * (o >> 5)], 1u << (o & 31u));  // rows are sorted: always true
            if constexpr (W) bval[1 + tid] = (double)(pre ? st_w[tid] : a.vw[x]);
          }
        } else {
          for (int i = tid; i < du; i += BLOCK) {
            const int32_t y = a.ind[ou + i];
            const uint32_t o = (uint32_t)(y - lo0);
            if (o < words * 32u) atomicOr(&bits[WS * (o >> 5)], 1u << (o & 31u));  // rows are sorted: always true
            if constexpr (W) bval[1 + i] = (double)a.vw[y];  // the i-th key in id order has rank i
          }
        }
        if (W && tid == 0) bval[0] = 0.0;
        __syncthreads();
        if constexpr (W) {
          const uint32_t per = (words + BLOCK - 1) / BLOCK;
          const uint32_t w0 = min(words, (uint32_t)tid * per), w1 = min(words, w0 + per);
          int c = 0;
          for (uint32_t i = w0; i < w1; ++i) c += __popc(bits[2 * i]);
          int r = block_excl_scan(c, s_wsum);
          for (uint32_t i = w0; i < w1; ++i) {
            bits[2 * i + 1] = (uint32_t)r + 1u;
            r += __popc(bits[2 * i]);
          }
          __syncthreads();
        }
        Chunk<W> cx;
        cx.u = u;
        cx.du = du;
        cx.su = row_sum_of<MODE>(a, u);
        cx.lo = lo0;
        cx.hi = hi0;
        cx.c = 0;
        cx.nch = 1;
        cx.kbase = ka;
        cx.cursor = cursor;
        cx.pcnt = pcnt;
        cx.pacc = pacc;
        cx.bits = bits;
        cx.bval = bval;
        cx.wzero = words;
        if (w == 0) {
          // lookahead (wave 0): reserve the next task and stage it; the loads
          // of all its entries are issued together (three round trips)
          int tn = 0;
          if (lane == 0) tn = atomicAdd(counter, 1);
          tn = __shfl(tn, 0);
          if (tn < ntasks) {
            const int2 tk2 = a.tasks[tbase + tn];
            const int ou2 = a.off[tk2.x];
            const int du2 = a.off[tk2.x + 1] - ou2;
            const int ka2 = ou2 + tk2.y * kTaskSpan;
            const int span2 = min(kTaskSpan, ou2 + du2 - ka2);
            const int32_t lo2 = a.ind[ou2], hi2 = a.ind[ou2 + du2 - 1];
            // loads are unconditional, their index clamped to the task's last
            // entry (span2 >= 1), so none waits behind a branch; only the
            // stores are limited to the task
            constexpr int NI = (kTaskSpan + kWave - 1) / kWave;
            int32_t xs[NI];
#pragma unroll
            for (int k = 0; k < NI; ++k) xs[k] = a.ind[ka2 + min(lane + k * kWave, span2 - 1)];
            int o0[NI], o1[NI];
#pragma unroll
            for (int k = 0; k < NI; ++k) {
              o0[k] = a.off[xs[k]];
              o1[k] = a.off[xs[k] + 1];
            }
            if constexpr (W) {
              if (du2 <= kTaskSpan) {
                float ws[NI];
#pragma unroll
                for (int k = 0; k < NI; ++k) ws[k] = a.vw[xs[k]];
#pragma unroll
                for (int k = 0; k < NI; ++k)
                  if (lane + k * kWave < span2) st_w[lane + k * kWave] = ws[k];
              }
            }
#pragma unroll
            for (int k = 0; k < NI; ++k) {
              const int i = lane + k * kWave;
              if (i < span2) {
                st_x[i] = xs[k];
                st_ox[i] = o0[k];
                st_dx[i] = o1[k] - o0[k];
              }
            }
            if (lane == 0) {
              s_hdr[0] = tk2.x;
              s_hdr[1] = ou2;
              s_hdr[2] = du2;
              s_hdr[3] = ka2;
              s_hdr[4] = lo2;
              s_hdr[5] = hi2;
              s_hdr[6] = 1;
            }
          }
          if (lane == 0) {
            s_task = tn;
            s_pre = tn < ntasks ? 1 : 0;
          }
        }
        for (;;) {
          int un = 0;
          if (lane == 0) un = atomicAdd(&s_next, 1);
          un = __shfl(un, 0);
          if (un >= nunits) break;
          int k = 0;
          while (un >= s_ubase[k + 1]) ++k;
          const int cls_k = kClasses - 1 - k;
          const int per = kWave / class_g(cls_k);
          const int first = s_cbase[k] + (un - s_ubase[k]) * per;
          run_unit_lds<MODE, W, false, kTkBitmap>(a, cx, cls_k, first, s_cbase[k + 1], ix, iox, idx_, ij, fast);
        }
        __syncthreads();
        for (int i = tid; i < s_cbase[kClasses]; i += BLOCK)
          emit_deferred<MODE, W>(a, u, du, cx.su, ka, i, ix[i], idx_[i], ij[i], cursor, pcnt, pacc);
        continue;
      }
      if constexpr (KIND != 1) {
        constexpr int TK = kQ ? kTkQuot : kTkHash;
        for (int c = 0; c < nch; ++c) {
          const int cs = ou + c * chk;
          const int ce = min(cs + chk, ou + du);
          const int clen = ce - cs;
          // a single-chunk table stays in the LDS for the task's later pieces
          const bool fresh = nch > 1 || pc == 0;
          Chunk<W> cx;
          if constexpr (kQ) {
            const int lb = qtable_log_buckets(clen, __builtin_ctz(Cfg::QNB));
            const int maxd = (1 << (lb - 8)) - 2;  // 2^(16 - rb) - 2
            cx.qtab = QTable<W>{qword, qval, lb, (uint32_t)((1 << lb) - 1), 24 - lb,
                                a.quot_max_disp > 0 ? min(maxd, a.quot_max_disp - 1) : maxd};
            if (fresh) {
              for (int i = tid; i < (1 << lb); i += BLOCK) cx.qtab.clear(i);
              if (tid == 0) {
                s_next = 0;
                s_fail = 0;
              }
              __syncthreads();
              build_qtable<W, BLOCK, KPT>(cx.qtab, a.ind, a.vw, cs, clen, tid, &s_fail, s_wsum);
              __syncthreads();
              if (s_fail) {
                // nothing of this task has been written yet (a failed build is
                // the first of its pieces): the 32-bit table redoes it
                if (tid == 0) a.redo[atomicAdd(a.num_redo, 1)] = t;
                built = false;
                break;
              }
            }
          } else {
            const int lb = block_table_log_buckets<W>(clen, __builtin_ctz(Cfg::NB));
            cx.tab = Table<W>{tkey, tval, lb, (uint32_t)((1 << lb) - 1), Cfg::NB * 4};
            if (fresh) {
              for (int i = tid; i < (4 << lb); i += BLOCK) cx.tab.clear(i);
              if (tid == 0) {
                s_next = 0;
                cx.tab.set_zero();
              }
              __syncthreads();
              build_table<W>(cx.tab, a.ind, a.vw, cs, clen, tid, BLOCK);
              __syncthreads();
            }
          }
          if (!fresh) {
            if (tid == 0) s_next = 0;
            __syncthreads();  // the piece's item list is complete
          }
          cx.u = u;
          cx.du = du;
          cx.su = row_sum_of<MODE>(a, u);
          cx.lo = a.ind[cs];
          cx.hi = a.ind[ce - 1];
          cx.c = c;
          cx.nch = nch;
          cx.kbase = ka;
          cx.cursor = cursor;
          cx.pcnt = pcnt;
          cx.pacc = pacc;
          if constexpr (kQ) {
            if (c == 0 && w == 0) {
              // lookahead (wave 0, then it joins the units): the next piece of
              // this task, or at its last piece the next task's header and
              // first piece; loads unconditional with indices clamped to the
              // piece (span >= 1), stores limited to it
              const bool same = pc + 1 < npieces;
              int tn = t, u2 = u, ou2 = ou, du2 = du, ka2 = ka + kTaskSpan, np2 = npieces;
              if (!same) {
                int r = 0;
                if (lane == 0) r = atomicAdd(counter, 1);
                tn = __shfl(r, 0);
                if (tn < ntasks) {
                  const int2 tk2 = a.tasks[tn];
                  u2 = tk2.x;
                  ou2 = a.off[u2];
                  du2 = a.off[u2 + 1] - ou2;
                  ka2 = ou2 + max(tk2.y, 0) * kTaskSpan;
                  np2 = tk2.y < 0 ? (du2 + kTaskSpan - 1) / kTaskSpan : 1;
                }
              }
              const bool staged = same || tn < ntasks;
              if (staged) {
                // in groups of four entries per lane: twelve live registers,
                // so the kernel's allocation (and the probe loops') does not
                // grow into spills
                const int span2 = min(kTaskSpan, ou2 + du2 - ka2);
                constexpr int NI = (kTaskSpan + kWave - 1) / kWave;
                constexpr int GI = 4;
                for (int k0 = 0; k0 < NI; k0 += GI) {
                  if (k0 * kWave >= span2) break;  // wave-uniform
                  int32_t xs[GI];
                  int o0[GI], o1[GI];
#pragma unroll
                  for (int k = 0; k < GI; ++k) xs[k] = a.ind[ka2 + min(lane + (k0 + k) * kWave, span2 - 1)];
#pragma unroll
                  for (int k = 0; k < GI; ++k) {
                    o0[k] = a.off[xs[k]];
                    o1[k] = a.off[xs[k] + 1];
                  }
#pragma unroll
                  for (int k = 0; k < GI; ++k) {
                    const int i = lane + (k0 + k) * kWave;
                    if (i < span2) {
                      st_x[i] = xs[k];
                      st_ox[i] = o0[k];
                      st_dx[i] = o1[k] - o0[k];
                    }
                  }
                }
              }
              if (lane == 0) {
                if (!same) {
                  s_task = tn;
                  s_rsv = 1;
                  s_hdr[0] = u2;
                  s_hdr[1] = ou2;
                  s_hdr[2] = du2;
                  s_hdr[3] = ka2;
                  s_hdr[4] = 0;
                  s_hdr[5] = 0;
                  s_hdr[6] = np2;
                }
                s_pre = staged ? 1 : 0;
              }
            }
          }
          for (;;) {
            int un = 0;
            if (lane == 0) un = atomicAdd(&s_next, 1);
            un = __shfl(un, 0);
            if (un >= nunits) break;
            int k = 0;
            while (un >= s_ubase[k + 1]) ++k;  // class slot k (0 = heaviest)
            const int cls_k = kClasses - 1 - k;
            const int per = kWave / class_g(cls_k);
            const int first = s_cbase[k] + (un - s_ubase[k]) * per;
            const int cend = s_cbase[k + 1];
            if (nch == 1)
              run_unit_lds<MODE, W, false, TK>(a, cx, cls_k, first, cend, ix, iox, idx_, ij, fast);
            else
              run_unit_lds<MODE, W, true, TK>(a, cx, cls_k, first, cend, ix, iox, idx_, ij, fast);
          }
          __syncthreads();
        }
        if (built) {
          const double su = row_sum_of<MODE>(a, u);
          for (int i = tid; i < s_cbase[kClasses]; i += BLOCK)
            emit_deferred<MODE, W>(a, u, du, su, ka, i, ix[i], idx_[i], ij[i], cursor, pcnt, pacc);
        }
      }
    }  // pieces
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
  // each lane tests four consecutive positions per step (one 16-byte load
  // for 4-byte flags, two for 8-byte ones; element loads when the array is
  // not 16-byte aligned, as a caller's result view may be)
  constexpr int V = 4;
  const bool vec = (reinterpret_cast<uintptr_t>(flag_arr) & 15) == 0;
  for (int64_t base = gw * kWave * V; base < a.m; base += nwaves * kWave * V) {
    const int64_t k0 = base + (int64_t)lane_id() * V;
    bool sb[V];
    if (vec && k0 + V <= a.m) {
      if constexpr (F64) {
        const ulonglong2* p2 = reinterpret_cast<const ulonglong2*>(reinterpret_cast<const uint64_t*>(flag_arr) + k0);
        const ulonglong2 x0 = p2[0], x1 = p2[1];
        sb[0] = x0.x == kSentinel64;
        sb[1] = x0.y == kSentinel64;
        sb[2] = x1.x == kSentinel64;
        sb[3] = x1.y == kSentinel64;
      } else {
        const uint4 x = *reinterpret_cast<const uint4*>(reinterpret_cast<const uint32_t*>(flag_arr) + k0);
        sb[0] = x.x == kSentinel32;
        sb[1] = x.y == kSentinel32;
        sb[2] = x.z == kSentinel32;
        sb[3] = x.w == kSentinel32;
      }
    } else {
#pragma unroll
      for (int j = 0; j < V; ++j) sb[j] = k0 + j < a.m && is_sentinel<F64>(flag_arr, k0 + j);
    }
    uint64_t badj[V];
#pragma unroll
    for (int j = 0; j < V; ++j) badj[j] = ballot(sb[j]);
    for (int j = 0; j < V; ++j) {
      uint64_t bad = badj[j];
      while (bad) {
        const int l = __ffsll((unsigned long long)bad) - 1;
        bad &= bad - 1;
        const int64_t kk = base + (int64_t)l * V + j;
        const int32_t u = wave_lower_bound(a.off, 0, a.n + 1, (int32_t)(kk + 1)) - 1;
        const int32_t v = a.ind[kk];
        uint32_t cnt;
        DD acc;
        wave_pair<W>(a.off, a.ind, a.vw, u, v, cnt, acc);
        if (lane_id() == 0) {
          const int du = a.off[u + 1] - a.off[u];
          const int dv = a.off[v + 1] - a.off[v];
          emit<MODE, W>(a, u, du, v, dv, kk, -1, cnt, acc, row_sum_of<MODE>(a, u), row_sum_of<MODE>(a, v));
        }
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

// Two task lists, one per table kind (0: hash table, 1: bitmap), each in
// degree-descending buckets; list 0 first in the task array.
struct Sched {
  int bucket_cnt[2][kBuckets];   // tasks per bucket
  int bucket_cur[2][kBuckets];   // emission cursors
  int num_tasks[2];
  int task_counter[3];           // hash-table list, bitmap list, redo list
  int num_small;
  int small_cur;
  int num_redo;                  // hash-table tasks the quotient table could not hold
};


__device__ __forceinline__ int bucket_of(int d) { return 31 - __clz(d); }

// Virtual row v -> graph row: identity without segments, else segment lookup
// (segs[i].y = rows before segment i, segs[nseg].y = total).
__device__ __forceinline__ int32_t seg_row(const int2* __restrict__ segs, int nseg, int32_t v) {
  if (nseg == 0) return v;
  int lo = 0, hi = nseg;  // segs[lo].y <= v < segs[hi].y
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (segs[mid].y <= v) lo = mid; else hi = mid;
  }
  return segs[lo].x + (v - segs[lo].y);
}

// Tasks of a row of d > small_max entries of table kind `kind`: a hash-table
// row whose keys fit one table chunk (d <= whole_max) is one task covering
// all its pieces of kTaskSpan entries, so its table is built once; other
// rows get one task per piece.
__device__ __forceinline__ int row_tasks(int d, int kind, int whole_max) {
  return (kind == 0 && d <= whole_max) ? 1 : (d + kTaskSpan - 1) / kTaskSpan;
}

template <bool W, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sched_count(int32_t rows, const int32_t* __restrict__ off,
                                                       const int32_t* __restrict__ ind, Sched* sc, int small_max,
                                                       size_t bm_budget, int whole_max, const int2* segs,
                                                       int nseg) {
  __shared__ int hist[2 * kBuckets + 1];
  for (int i = threadIdx.x; i < 2 * kBuckets + 1; i += BLOCK) hist[i] = 0;
  __syncthreads();
  for (int32_t v = blockIdx.x * BLOCK + threadIdx.x; v < rows; v += gridDim.x * BLOCK) {
    const int32_t u = seg_row(segs, nseg, v);
    const int o = off[u], d = off[u + 1] - o;
    if (d <= 0) continue;
    if (d <= small_max) {
      atomicAdd(&hist[2 * kBuckets], 1);
    } else {
      const int kind = bitmap_fits<W>(ind[o], ind[o + d - 1], d, bm_budget) ? 1 : 0;
      atomicAdd(&hist[kind * kBuckets + bucket_of(d)], row_tasks(d, kind, whole_max));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 2 * kBuckets; i += BLOCK)
    if (hist[i]) atomicAdd(&sc->bucket_cnt[i / kBuckets][i % kBuckets], hist[i]);
  if (threadIdx.x == 0 && hist[2 * kBuckets]) atomicAdd(&sc->num_small, hist[2 * kBuckets]);
}

static __global__ void k_sched_init(Sched* sc) {
  int* p = reinterpret_cast<int*>(sc);
  for (int i = threadIdx.x; i < (int)(sizeof(Sched) / sizeof(int)); i += blockDim.x) p[i] = 0;
}

static __global__ void k_sched_scan(Sched* sc) {
  if (threadIdx.x != 0) return;
  int acc = 0;
  for (int k = 0; k < 2; ++k) {
    const int first = acc;
    for (int b = kBuckets - 1; b >= 0; --b) {
      sc->bucket_cur[k][b] = acc;
      acc += sc->bucket_cnt[k][b];
    }
    sc->num_tasks[k] = acc - first;
    sc->task_counter[k] = 0;
  }
  sc->small_cur = 0;
}

// Each workgroup takes a contiguous range of rows: it counts the range's
// rows per bin (LDS atomics), reserves each bin's places with one global
// atomic, then walks the range again and writes. (Reserving per 256 rows put
// one atomic per bin and step on the same few counters -- the small-row
// cursor above all -- which serialise: 0.20 ms on RMAT-22.)
template <bool W, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sched_emit(int32_t rows, const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ ind, Sched* sc, int2* tasks,
                                                      int32_t* small_rows, int small_max, size_t bm_budget,
                                                      int whole_max, const int2* segs, int nseg) {
  constexpr int kSmallBin = 2 * kBuckets;
  __shared__ int hist[kSmallBin + 1];
  __shared__ int basep[kSmallBin + 1];
  const int64_t per = ((int64_t)rows + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = per * blockIdx.x, b1 = b0 + per;
  const int32_t r0 = (int32_t)(b0 < rows ? b0 : rows);
  const int32_t r1 = (int32_t)(b1 < rows ? b1 : rows);
  // bin and task count of graph row u (b = -1: empty row)
  auto classify = [&](int32_t u, int& d, int& b, int& cntv) {
    const int o = off[u];
    d = off[u + 1] - o;
    b = -1;
    cntv = 0;
    if (d <= 0) return;
    if (d <= small_max) {
      b = kSmallBin;
      cntv = 1;
    } else {
      const int kind = bitmap_fits<W>(ind[o], ind[o + d - 1], d, bm_budget) ? 1 : 0;
      b = kind * kBuckets + bucket_of(d);
      cntv = row_tasks(d, kind, whole_max);
    }
  };
  for (int i = threadIdx.x; i < kSmallBin + 1; i += BLOCK) hist[i] = 0;
  __syncthreads();
  for (int32_t v = r0 + threadIdx.x; v < r1; v += BLOCK) {
    int d, b, cntv;
    classify(seg_row(segs, nseg, v), d, b, cntv);
    if (b >= 0) atomicAdd(&hist[b], cntv);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kSmallBin + 1; i += BLOCK) {
    if (hist[i])
      basep[i] = i == kSmallBin ? atomicAdd(&sc->small_cur, hist[i])
                                : atomicAdd(&sc->bucket_cur[i / kBuckets][i % kBuckets], hist[i]);
    hist[i] = 0;  // now the range's running cursor per bin
  }
  __syncthreads();
  for (int32_t v = r0 + threadIdx.x; v < r1; v += BLOCK) {
    const int32_t u = seg_row(segs, nseg, v);
    int d, b, cntv;
    classify(u, d, b, cntv);
    if (b < 0) continue;
    const int at = basep[b] + atomicAdd(&hist[b], cntv);
    if (b == kSmallBin) {
      small_rows[at] = u;
    } else if (b < kBuckets && cntv == 1 && d > kTaskSpan) {
      tasks[at] = make_int2(u, -1);  // a whole-row task
    } else {
      for (int p = 0; p < cntv; ++p) tasks[at + p] = make_int2(u, p);  // piece tasks
    }
  }
}

// ---------------------------------------------------------------------------
// Host side.
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

inline int64_t max_tasks(int32_t n, int64_t m) {
  // each row with degree > small_max (>= 1) contributes ceil(d / kTaskSpan) <= d / kTaskSpan + 1
  // tasks, and there are at most min(n, m / 2) such rows
  return m / kTaskSpan + std::min<int64_t>(n, m / 2) + 1;
}

template <int MODE, bool W>
hipError_t run_engine_t(const EngineArgs& base, void* ws, bool fixup, hipStream_t stream) {
  constexpr int SB = 256;   // small-row kernel block
  constexpr int BB = JC_BLOCK;  // block kernel block
  constexpr int PB = 256;   // schedule kernels
  EngineArgs a = base;
  if (a.nseg > 0 && fixup) return hipErrorInvalidValue;  // segments need a symmetric graph
  const int32_t rows = a.nseg > 0 ? a.seg_rows : a.n;
  const int64_t edges = a.nseg > 0 ? a.seg_edges : a.m;
  if (a.small_max < 1 || a.small_max > kSmallMax) a.small_max = kSmallMax;
  if (a.small_blocks_per_cu < 1) a.small_blocks_per_cu = 16;
  // buffer-load offsets reach 4 * (m + 256) bytes
  if (a.m * 4 + 8192 > (int64_t)UINT32_MAX) a.wide_index = 1;
  a.wide_index = a.wide_index ? 1 : 0;
  unsigned char* p = static_cast<unsigned char*>(ws);
  Sched* sc = reinterpret_cast<Sched*>(p);
  p += align_up(sizeof(Sched), 256);
  int2* tasks = reinterpret_cast<int2*>(p);
  p += align_up((size_t)max_tasks(rows, edges) * sizeof(int2), 256);
  int32_t* small = reinterpret_cast<int32_t*>(p);
  p += align_up((size_t)std::max<int32_t>(rows, 1) * sizeof(int32_t), 256);
  // quotient tables for the hash-table tasks when n < kQuotMaxN
  const bool quot = JC_QTABLE && a.n < kQuotMaxN && a.hash_table != 1;
  a.redo = quot ? reinterpret_cast<int*>(p) : nullptr;
  a.num_redo = &sc->num_redo;
  a.redo_pass = 0;
  a.tasks = tasks;
  a.num_tasks = sc->num_tasks;
  a.task_counter = sc->task_counter;
  a.small_rows = small;
  a.num_small = &sc->num_small;

  const int cus = device_cu_count();
  hipError_t err;
  hipLaunchKernelGGL(k_sched_init, dim3(1), dim3(128), 0, stream, sc);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  if (rows > 0) {
    const int g = std::min<int64_t>((rows + PB - 1) / PB, (int64_t)cus * 8);
    using Cfg = BlockCfg<W>;
    // bitmap tasks take all of the LDS but the item list, the staging area
    // and the result slots with one workgroup per CU, or 32 KiB with two
    const int bm_per_cu = a.bitmap_blocks_per_cu == 2 ? 2 : 1;
    const size_t bm_rest = Cfg::bm_rest;
    const size_t bm_budget = bm_per_cu == 2 ? kBitmapLds2 : Cfg::bm_budget1;
    const size_t lds1 = bm_budget + bm_rest;
    a.bm_budget = (uint32_t)bm_budget;
    // hash-table rows up to one chunk of keys are whole-row tasks
    const int whole_max = quot ? Cfg::QCH : Cfg::CH;
    {
      ProfScope ps("sched", stream);
      hipLaunchKernelGGL((k_sched_count<W, PB>), dim3(g), dim3(PB), 0, stream, rows, a.off, a.ind, sc, a.small_max,
                         bm_budget, whole_max, a.segs, a.nseg);
      hipLaunchKernelGGL(k_sched_scan, dim3(1), dim3(64), 0, stream, sc);
      hipLaunchKernelGGL((k_sched_emit<W, PB>), dim3(g), dim3(PB), 0, stream, rows, a.off, a.ind, sc, tasks, small,
                         a.small_max, bm_budget, whole_max, a.segs, a.nseg);
      if ((err = hipGetLastError()) != hipSuccess) return err;
    }

    // hash-table tasks, then bitmap tasks: one persistent grid each (a grid
    // whose list is empty returns at once)
    // (quotient-table kernel first when n allows it, then the 32-bit kernel
    // over the tasks it could not hold, usually none)
    int occ0 = 0, occ1 = 0, occ2 = 0;
    err = hipFuncSetAttribute(reinterpret_cast<const void*>(k_owner_block<MODE, W, BB, 0>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)Cfg::lds_bytes);
    if (err != hipSuccess) return err;
    err = hipFuncSetAttribute(reinterpret_cast<const void*>(k_owner_block<MODE, W, BB, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
    if (err != hipSuccess) return err;
    err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ0, k_owner_block<MODE, W, BB, 0>, BB, Cfg::lds_bytes);
    if (err != hipSuccess) return err;
    err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ1, k_owner_block<MODE, W, BB, 1>, BB, lds1);
    if (err != hipSuccess) return err;
    if (quot) {
      err = hipFuncSetAttribute(reinterpret_cast<const void*>(k_owner_block<MODE, W, BB, 2>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)Cfg::q_lds_bytes);
      if (err != hipSuccess) return err;
      err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ2, k_owner_block<MODE, W, BB, 2>, BB,
                                                         Cfg::q_lds_bytes);
      if (err != hipSuccess) return err;
      if (occ2 < 1) return hipErrorInvalidConfiguration;
    }
    if (occ0 < 1 || occ1 < 1) return hipErrorInvalidConfiguration;
    occ1 = std::min(occ1, bm_per_cu);
    {
      ProfScope ps("owner_block", stream);
      if (quot) {
        hipLaunchKernelGGL((k_owner_block<MODE, W, BB, 2>), dim3(cus * occ2), dim3(BB), Cfg::q_lds_bytes, stream,
                           a);
        EngineArgs ar = a;
        ar.redo_pass = 1;
        hipLaunchKernelGGL((k_owner_block<MODE, W, BB, 0>), dim3(cus * occ0), dim3(BB), Cfg::lds_bytes, stream,
                           ar);
      } else {
        hipLaunchKernelGGL((k_owner_block<MODE, W, BB, 0>), dim3(cus * occ0), dim3(BB), Cfg::lds_bytes, stream, a);
      }
      hipLaunchKernelGGL((k_owner_block<MODE, W, BB, 1>), dim3(cus * occ1), dim3(BB), lds1, stream, a);
      if ((err = hipGetLastError()) != hipSuccess) return err;
    }
    const int sg = std::min<int64_t>((rows + (SB / kWave) - 1) / (SB / kWave), (int64_t)cus * a.small_blocks_per_cu);
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

}  // namespace jc
