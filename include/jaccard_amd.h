/*
 * jaccard_amd.h -- C ABI of the MI355X (gfx950) Jaccard library.
 *
 * Edge-connected Jaccard similarity over an int32 CSR graph, the hot path of
 * the SYCL reference (vtsynergy/SYCL-Jaccard). Every entry point below names
 * the reference interface it replaces (paths relative to the reference root).
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (hipMalloc'd or equivalent).
 *     The caller owns inputs and outputs, as in the reference ("memory needs
 *     to be pre-allocated by caller", standalone_algorithms.hpp:64-65).
 *   - Calls are ordered on `stream` (NULL = the default stream) and return
 *     without waiting for the device, except where marked "synchronous".
 *     Temporaries are stream-ordered allocations freed on the same stream,
 *     unless a workspace is supplied (jc_jaccard_ex).
 *   - Return value: JC_OK (0) or a negative JC_ERR_* code; jc_strerror()
 *     describes it. HIP runtime failures return JC_ERR_HIP and
 *     jc_last_hip_error() gives the hipError_t of the failing call on this
 *     thread. (The reference returns void and prints SYCL exceptions,
 *     jaccard.cpp:178-180.)
 *   - CSR preconditions (the reference's CSR builder guarantees them,
 *     readMtxToCSR.cpp:152-208): offsets[0] = 0, non-decreasing,
 *     offsets[n] = m; 0 <= indices[k] < n; columns strictly increasing inside
 *     each row. jc_validate_csr() checks them. A symmetric CSR (every (u,v)
 *     also stored as (v,u), main.cpp:98-112) takes the fast path; any other
 *     CSR is still computed exactly, with a fix-up pass for unmatched edges.
 *   - Unweighted results are bit-identical to the reference arithmetic
 *     I / ((float)(d_u + d_v) - I). Weighted results (vertex weights W):
 *     J = I_W / (S(u) + S(v) - I_W), S(u) = sum_{k in N(u)} W[k],
 *     I_W = sum_{k in N(u) & N(v)} W[k], accumulated exactly (double-double,
 *     or plain fp64 where that is exact, jc_weighted_sum_mode) and divided in
 *     fp64, rounded once to fp32.
 */
#ifndef JACCARD_AMD_H
#define JACCARD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hipStream_t (struct ihipStream_t *): pass a hipStream_t directly. */
struct ihipStream_t;
typedef struct ihipStream_t *jc_stream_t;

enum {
  JC_OK = 0,
  JC_ERR_INVALID_ARG = -1,    /* null pointer, negative size, bad flag */
  JC_ERR_INVALID_CSR = -2,    /* jc_validate_csr found a violated precondition */
  JC_ERR_OUT_OF_MEMORY = -3,  /* device allocation failed */
  JC_ERR_HIP = -4,            /* a HIP runtime call failed, see jc_last_hip_error() */
  JC_ERR_WORKSPACE = -5,      /* supplied workspace smaller than required */
  JC_ERR_SCRATCH_LIMIT = -6   /* jc_jaccard_noalloc: the graph needs more than
                                 JC_NOALLOC_SCRATCH_BYTES of workspace */
};

/* Graph in CSR form: GraphCSRView<int32_t,int32_t,float>
 * (standalone_csr.hpp:44-105; edge_data, number_of_vertices, number_of_edges,
 * offsets, indices). All pointers are device pointers. */
typedef struct jc_csr {
  int32_t n;               /* number_of_vertices */
  int32_t m;               /* number_of_edges (stored, directed) */
  const int32_t *offsets;  /* n + 1 */
  const int32_t *indices;  /* m */
  const float *edge_w;     /* m edge weights, or NULL (used by jc_vertex_weights) */
} jc_csr;

#define JC_EDGE_CENTRIC 1
#define JC_VERTEX_CENTRIC 0

/* Skip the sentinel fill and fix-up pass: the caller asserts the CSR is
 * symmetric (for instance it was checked once with jc_csr_is_symmetric). */
#define JC_FLAG_ASSUME_SYMMETRIC 1u

/* ---- Algorithm level -------------------------------------------------- */

/* sygraph::jaccard<edge_centric, int, int, float>(graph, [weights,] result, q)
 *   declarations standalone_algorithms.hpp:67-69 (weighted), :86-87
 *   (unweighted); definitions jaccard.cpp:926-955; pipeline jaccard.cpp:784-899.
 * vertex_weights: n vertex weights (weighted Jaccard), or NULL (unweighted).
 * result: m outputs, one per stored edge in CSR order.
 * edge_centric: JC_EDGE_CENTRIC (fused kernel) or JC_VERTEX_CENTRIC
 * (row sum -> intersection -> weight kernels). Both give the same bits.
 * Differences from the reference: the reference's edge-centric path ignores
 * vertex weights (jaccard.cpp:790-819 always runs the unweighted kernel);
 * here edge-centric honours them. The EC "non zero pairs" line
 * (jaccard.cpp:811-819) is jc_nonzero_count, not a side effect. */
int jc_jaccard(const jc_csr *g, const float *vertex_weights, float *result, int edge_centric,
               jc_stream_t stream);

/* jc_jaccard with flags (JC_FLAG_*) and an optional caller workspace of at
 * least jc_jaccard_workspace_size() bytes (NULL: allocate per call). With a
 * workspace the call allocates nothing. No reference counterpart: the
 * reference allocates temporaries inside the call (jaccard.cpp:931-934). */
int jc_jaccard_ex(const jc_csr *g, const float *vertex_weights, float *result, int edge_centric,
                  unsigned flags, void *workspace, size_t workspace_bytes, jc_stream_t stream);
size_t jc_jaccard_workspace_size(int32_t n, int32_t m, int weighted, int edge_centric);

/* Launch tuning for jc_jaccard_tuned. None of these change results; zero
 * fields take the defaults. No reference counterpart (the reference's launch
 * shapes are fixed, jaccard.cpp:264-276, 403-418, 745-753). */
typedef struct jc_tuning {
  int32_t small_max;           /* degree bin threshold: rows up to this degree run one
                                  wave per row, longer rows one workgroup per task
                                  (1..128, default 128) */
  int32_t small_blocks_per_cu; /* grid of the one-wave-per-row kernel (default 16) */
  int32_t wide_index;          /* nonzero: stream rows with 64-bit addresses even when
                                  the index array is small enough for 32-bit buffer
                                  offsets (arrays of 2^30 entries or more always use
                                  them); for testing both load paths */
  int32_t bitmap_blocks_per_cu; /* workgroups per CU asked for the bitmap-table task
                                  kernel (1 or 2, default 1). With 2 its LDS table
                                  region shrinks to 32 KiB (rows whose bitmap no longer
                                  fits use the hash-table kernel); the grid is the CU
                                  count times what the kernel's registers and LDS let
                                  reside, at most this value */
  int32_t hash_table;          /* hash-table tasks: 0 (default) quotient tables (16-bit
                                  keys) when n < 2^24, else 32-bit tables; 1 always
                                  32-bit tables */
  int32_t quot_max_disp;       /* for tests: 0 (default) or k >= 1, which caps how far
                                  (k - 1 buckets) a quotient-table key may move from its
                                  home bucket; a task whose table then does not build is
                                  redone with the 32-bit table (1 forces that for every
                                  chunk with a full bucket) */
  int32_t reserved[2];
} jc_tuning;

/* jc_jaccard_ex that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid allocation). No reference counterpart. Its workspace is a
 * fixed array of JC_NOALLOC_SCRATCH_BYTES that the library holds on every
 * device it runs on until the process ends; a graph whose
 * jc_jaccard_workspace_size() exceeds it gives JC_ERR_SCRATCH_LIMIT (32 MiB:
 * about 1.18M vertices edge-centric weighted, 2.05M unweighted, at 16 to 30
 * stored edges per vertex; vertex-centric also needs 8 or 16 bytes per
 * stored edge). Inputs and the result are the
 * caller's; the result is the bits of jc_jaccard_ex. All calls on one device
 * share that device's scratch: calls that overlap on one device (other
 * streams or threads) race and must be serialized by the caller. */
#define JC_NOALLOC_SCRATCH_BYTES (32u << 20)
int jc_jaccard_noalloc(const jc_csr *g, const float *vertex_weights, float *result, int edge_centric,
                       unsigned flags, jc_stream_t stream);

/* jc_jaccard_ex with launch tuning (NULL = defaults). */
int jc_jaccard_tuned(const jc_csr *g, const float *vertex_weights, float *result, int edge_centric,
                     unsigned flags, void *workspace, size_t workspace_bytes, const jc_tuning *tuning,
                     jc_stream_t stream);

/* Automatic configuration from graph properties, the counterpart of the
 * reference driver's "TODO: Automatic selection based on graph properties"
 * (main.cpp:203), which there picks nothing (one of the FORCE variables must
 * be set, main.cpp:48-61).
 *
 * jc_degree_stats (synchronous) gathers on the device: rows[b] and
 * entries[b], the rows of degree in [2^b, 2^(b+1)) and their stored entries;
 * isolated rows; long_rows, the rows above the one-wave-per-row bin (degree
 * > 128); long_bitmap_rows[w], those among them that the workgroup kernel
 * gives a bitmap table (w = 0 unweighted, 1 weighted layout, one workgroup
 * per CU): a row with first and last column lo, hi and degree d gets one when
 * W = (hi - lo) / 32 + 2 words of 4 bytes (weighted: W words of 8 bytes plus
 * 8 * (d + 1) bytes of weights) fit bitmap_table_bytes[w] and, weighted, the
 * 8 * W bytes of words are at most 32 KiB; the largest degree. */
typedef struct jc_graph_stats {
  int64_t rows[32];
  int64_t entries[32];
  int64_t isolated;
  int64_t long_rows;
  int64_t long_bitmap_rows[2];
  int32_t max_degree;
  int32_t reserved;
  int64_t bitmap_table_bytes[2];
} jc_graph_stats;
int jc_degree_stats(const jc_csr *g, jc_graph_stats *stats, jc_stream_t stream);

/* The algorithm and launch tuning for a graph with statistics *stats and n
 * vertices (host only, no device work; *edge_centric and *tuning are
 * outputs). The rule:
 *   - edge-centric always: both algorithms give the same values from the same
 *     sums, and the vertex-centric one also writes and reads two m-length
 *     arrays (weight_i, weight_s) and runs the jw pass;
 *   - small_max 128: every row up to the largest the one-wave-per-row
 *     kernel's table holds goes to that kernel (smaller bins were slower in
 *     every family measured, up to 1.4x);
 *   - small_blocks_per_cu 32 (best or within 1% in every family);
 *   - bitmap_blocks_per_cu 1 (two workgroups per CU shrink the bitmap table
 *     and were up to 2.4x slower);
 *   - hash_table 0 (quotient tables) below 2^24 vertices, 1 (32-bit tables)
 *     from there on.
 * These come from a sweep of every jc_tuning field over thirteen graph
 * families (RMAT 2^14 to 2^22, weighted and not, Chung-Lu 2^18 and 2^20, the
 * two halves of a 4096-graph batch), where the chosen configuration was the
 * fastest or within 1% of it on eleven; no field changes results. The
 * table kind of each row (bitmap or hash) is decided per row from its
 * degree and column span by the scheduler, as jc_degree_stats reports. */
int jc_config_from_stats(const jc_graph_stats *stats, int32_t n, int weighted, int *edge_centric,
                         jc_tuning *tuning);

/* jc_degree_stats then jc_config_from_stats (synchronous). */
int jc_select_config(const jc_csr *g, int weighted, int *edge_centric, jc_tuning *tuning, jc_stream_t stream);

/* sygraph::jaccard_list<int, int, float>(graph, [weights,] num_pairs, first, second, result, q)
 *   declarations standalone_algorithms.hpp:109-112, :131-133; definitions
 *   jaccard.cpp:957-988 via jaccard_pairs jaccard.cpp:901-923.
 * result[i] = J(first[i], second[i]) for any vertex pair (not only edges).
 * Ids outside [0, n) give NaN. Same arithmetic as jc_jaccard, so the value of
 * a stored edge matches the full-graph result bit for bit. The reference's
 * version is compiled out (Makefile:33 -DDISABLE_LIST) and indexes weight_i
 * by the wrong variable (jaccard.cpp:517-531); this implements the intended
 * per-pair semantics. */
int jc_jaccard_list(const jc_csr *g, const float *vertex_weights, int32_t num_pairs,
                    const int32_t *first, const int32_t *second, float *result,
                    jc_stream_t stream);

/* Weighted-VC preamble of the reference driver, main.cpp:211-275:
 * W[k] = sum of the edge weights of row k (g->edge_w, required). */
int jc_vertex_weights(const jc_csr *g, float *W, jc_stream_t stream);

/* jc_vertex_weights with a caller workspace of at least
 * jc_vertex_weights_workspace_size(n) bytes: allocates nothing (for stream
 * capture). No reference counterpart. */
int jc_vertex_weights_ex(const jc_csr *g, float *W, void *workspace, size_t workspace_bytes,
                         jc_stream_t stream);
size_t jc_vertex_weights_workspace_size(int32_t n);

/* "non zero pairs" count of the EC path, jaccard.cpp:811-819: the number of
 * result entries greater than thresh (the reference uses 1e-5f).
 * Synchronous: waits for the stream, writes *count on the host. */
int jc_nonzero_count(int32_t m, const float *result, float thresh, int64_t *count,
                     jc_stream_t stream);

/* ---- Kernel level (the reference's Kernel::invoke entry points) -------- */

/* FillKernel<float>::invoke, jaccard.hpp:51-52 / jaccard.cpp:161-182. */
int jc_fill(int64_t n, float *x, float value, jc_stream_t stream);

/* Jaccard_RowSumKernel<weighted>::invoke, jaccard.hpp:104-107 /
 * jaccard.cpp:263-313: work[u] = (float)degree(u) when v is NULL, else
 * sum_{k in row u} v[indices[k]]. Asynchronous (the reference waits, :308). */
int jc_rowsum(int32_t n, int32_t m, const int32_t *offsets, const int32_t *indices,
              const float *v, float *work, jc_stream_t stream);

/* Jaccard_IsKernel<weighted>::invoke, jaccard.hpp:151-156 / jaccard.cpp:402-458:
 * weight_s[j] = work[u] + work[v]; weight_i[j] += I(u, v) (v NULL: count,
 * else sum of v over the intersection). Like the reference it adds into
 * weight_i, so zero it first (jc_fill), as jaccard.cpp:836 does. */
int jc_intersect(int32_t n, int32_t m, const int32_t *offsets, const int32_t *indices,
                 const float *v, const float *work, float *weight_i, float *weight_s,
                 jc_stream_t stream);

/* Jaccard_IsPairsKernel<weighted>::invoke, jaccard.hpp:210-217 /
 * jaccard.cpp:542-605, per pair i: weight_s[i] = work[first[i]] + work[second[i]];
 * weight_i[i] += I(first[i], second[i]). */
int jc_intersect_pairs(int32_t n, int32_t num_pairs, const int32_t *offsets,
                       const int32_t *indices, const int32_t *first, const int32_t *second,
                       const float *v, const float *work, float *weight_i, float *weight_s,
                       jc_stream_t stream);

/* Jaccard_JwKernel::invoke, jaccard.hpp:235-237 / jaccard.cpp:622-653:
 * weight_j[j] = weight_i[j] / (weight_s[j] - weight_i[j]). */
int jc_jw(int32_t m, const float *weight_i, const float *weight_s, float *weight_j,
          jc_stream_t stream);

/* Jaccard_ec_scan::invoke, jaccard.hpp:256-258 / jaccard.cpp:666-700:
 * dest_ind[k] = row of edge k, weight_j[k] = 0. */
int jc_ec_scan(int32_t m, int32_t n, const int32_t *offsets, int32_t *dest_ind, float *weight_j,
               jc_stream_t stream);

/* Jaccard_ec_unweighted::invoke, jaccard.hpp:280-283 / jaccard.cpp:745-782:
 * weight_j[k] = I / ((float)(d_u + d_v) - I) for every stored edge. The
 * result overwrites weight_j; dest_ind (from jc_ec_scan) is accepted for
 * signature parity and not needed: rows are walked directly. */
int jc_ec_unweighted(int32_t m, int32_t n, const int32_t *offsets, const int32_t *indices,
                     const int32_t *dest_ind, float *weight_j, jc_stream_t stream);

/* ---- Graph ingest -------------------------------------------------------- */

/* The driver's graph preparation (main.cpp:98-120; readMtxToCSR.cpp:116-208,
 * there with std::set on the host), on the device: from num_edges directed
 * edges (src, dst, optional edge_w; device arrays) builds a CSR with columns
 * sorted, self loops dropped and repeated (src, dst) merged, adding the
 * reverse of every edge when symmetrize != 0. A merged pair keeps the weight
 * of its first occurrence, originals before reverses. offsets: n + 1 entries;
 * indices and weights_out (NULL: no weights): capacity entries, where
 * num_edges (2 * num_edges with symmetrize) always suffices; *m_out receives
 * the stored edge count. Ids outside [0, n) give JC_ERR_INVALID_ARG, a
 * capacity below the count JC_ERR_WORKSPACE. Synchronous. */
int jc_csr_from_edges(int32_t n, int64_t num_edges, const int32_t *src, const int32_t *dst,
                      const float *edge_w, int symmetrize, int32_t *offsets, int32_t *indices,
                      float *weights_out, int64_t capacity, int64_t *m_out, jc_stream_t stream);

/* ---- Utilities (no reference counterpart) ------------------------------ */

/* Synchronous CSR check: JC_OK or JC_ERR_INVALID_CSR. *what (optional)
 * receives a bit mask: 1 offsets, 2 column range, 4 row order. */
int jc_validate_csr(const jc_csr *g, int *what, jc_stream_t stream);

/* Synchronous symmetry check: *symmetric = 1 if every stored (u,v) has (v,u). */
int jc_csr_is_symmetric(const jc_csr *g, int *symmetric, jc_stream_t stream);

/* How the weighted entry points (jc_jaccard, jc_jaccard_list, the batched and
 * served forms) add vertex weights W on graph g: JC_SUM_FP64 when every W is
 * a non-negative multiple of 2^e and every row sum is below 2^(e + 53), so
 * plain fp64 sums are exact; else JC_SUM_DOUBLE_DOUBLE (exact for spans up
 * to 104 bits). Both give the same bits wherever both are exact. The decision
 * is the one those calls make on the device for the same g and W.
 * Synchronous. No reference counterpart (the reference adds in fp32,
 * jaccard.cpp:53-137, 370-396). */
#define JC_SUM_FP64 0
#define JC_SUM_DOUBLE_DOUBLE 1
int jc_weighted_sum_mode(const jc_csr *g, const float *vertex_weights, int *mode, jc_stream_t stream);

/* Per-kernel device timing, the counterpart of the reference's EVENT_PROFILE
 * build (wait_and_print, standalone_algorithms.hpp:37-47, used at
 * jaccard.cpp:798-809, 875-896): while enabled, each kernel launch of the
 * library is bracketed by hipEvents on its stream. jc_profile_query
 * (synchronous) sums the launches of one kernel name ("owner_block",
 * "owner_small", "sched", "fixup", "fill", "rowsum", "jw", "vertex_weights",
 * "ec_scan", "pairs", "count_above"; NULL = all). */
int jc_profile_enable(int on);
int jc_profile_reset(void);
int jc_profile_query(const char *kernel, int *count, double *total_ms);
int jc_profile_names(char *buf, size_t len);

/* ---- Serving (no reference counterpart: the reference computes one graph
 * per call, jaccard.cpp:926-955) ----------------------------------------
 *
 * A server keeps a set of graphs resident in HBM as one arena (their disjoint
 * union with global ids) and replays a stream of timed requests against it:
 * whole-graph edge-centric Jaccard, or the Jaccard of k vertex pairs. In
 * JC_SERVE_CONTINUOUS mode requests that arrive while others run are admitted
 * into the next batch as soon as a slot (a HIP stream) and memory from the
 * HBM budget are free: one batch runs the whole-graph requests of several
 * graphs in one owner-engine launch over their row ranges and all pair
 * requests in one pair launch, and up to `slots` batches overlap on the
 * device. JC_SERVE_FIFO serves one request at a time in arrival order with
 * the plain entry points (jc_jaccard, jc_jaccard_list) on each graph's own
 * CSR, the per-request baseline. Both modes give the same bits for the same
 * request (exact accumulation; per-row reductions independent of batching). */
enum { JC_SERVE_CONTINUOUS = 0, JC_SERVE_FIFO = 1 };
enum { JC_REQ_FULL = 0, JC_REQ_PAIRS = 1 };

typedef struct jc_server_config {
  int32_t slots;              /* batches in flight at once, one stream each (1..16) */
  int32_t max_batch_requests; /* requests per batch (>= 1) */
  int64_t max_batch_edges;    /* stored edges of whole-graph work per batch; a
                                 single larger graph still runs alone */
  int64_t max_batch_pairs;    /* pairs per batch; a single larger request runs alone */
  size_t hbm_budget_bytes;    /* device memory for batch workspaces and pair staging;
                                 a request is admitted only when its batch fits */
  int32_t mode;               /* JC_SERVE_CONTINUOUS or JC_SERVE_FIFO */
  int32_t reserved[5];
} jc_server_config;

typedef struct jc_server jc_server;

/* Builds the resident arena from num_graphs symmetric CSR graphs (device
 * pointers, local ids; edge_w unused) and their vertex weights (device
 * arrays of n floats, or vertex_weights = NULL for unweighted Jaccard), and
 * allocates the HBM budget. Synchronous. The graphs must stay allocated for
 * the server's lifetime (JC_SERVE_FIFO reads them). A graph that is not
 * symmetric (some stored (u,v) without (v,u)) gives JC_ERR_INVALID_CSR:
 * batched passes compute each pair once for both stored directions. */
int jc_server_create(int32_t num_graphs, const jc_csr *graphs, const float *const *vertex_weights,
                     const jc_server_config *cfg, jc_server **out);
int jc_server_destroy(jc_server *s);
/* Bytes of device memory the arena holds (outside the budget). */
size_t jc_server_arena_bytes(const jc_server *s);

typedef struct jc_request {
  int32_t graph;         /* index into the graphs given to jc_server_create */
  int32_t kind;          /* JC_REQ_FULL or JC_REQ_PAIRS */
  int32_t num_pairs;     /* JC_REQ_PAIRS: k */
  int32_t reserved;
  const int32_t *first;  /* JC_REQ_PAIRS: HOST arrays of k local vertex ids (ids */
  const int32_t *second; /*   outside the graph give NaN) */
  float *result;         /* DEVICE output: m floats (FULL) or k floats (PAIRS) */
  double arrival;        /* seconds after the start of jc_server_run */
} jc_request;

typedef struct jc_request_stats {
  double admitted;       /* seconds after start: work enqueued */
  double completed;      /* seconds after start: result in `result` */
  int32_t batch;         /* batch number (JC_SERVE_FIFO: request order) */
  int32_t status;        /* JC_OK or a JC_ERR_* code for this request */
} jc_request_stats;

typedef struct jc_server_run_info {
  int32_t batches;           /* batches launched */
  int32_t max_in_flight;     /* most batches in flight at once */
  size_t peak_budget_bytes;  /* most budget memory held at once */
  double wall_seconds;       /* start to last completion */
} jc_server_run_info;

/* Replays `count` requests: each becomes visible at its arrival time (host
 * monotonic clock from the start of the call) and is served in the server's
 * mode. Returns when all are complete; stats[i] describes reqs[i]. A request
 * that can never fit the budget gets status JC_ERR_OUT_OF_MEMORY. Synchronous;
 * one call at a time per server. */
int jc_server_run(jc_server *s, const jc_request *reqs, int32_t count, jc_request_stats *stats,
                  jc_server_run_info *info);

const char *jc_strerror(int code);
int jc_last_hip_error(void);
const char *jc_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* JACCARD_AMD_H */
