/*
 * CPU oracle for the Jaccard hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates the reference's arithmetic in plain C so the HIP library
 * can be checked against it. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or, for
 * the baseline, as the thing timed on the host). The product never calls it.
 *
 * Functions and the reference code each one follows (paths relative to the
 * reference repository root):
 *
 *   or_ec_unweighted   jaccard.cpp:654-664 (row expansion) + jaccard.cpp:711-742
 *                      (per stored edge: ref = shorter row, cur = longer row,
 *                      tie -> ref = source row; binary search every ref element
 *                      in cur; count held in fp32; out = I / ((float)(Ni+Nj) - I))
 *   or_ec_edges        the same per-edge loop over an explicit edge subset
 *                      (the bounded CPU-baseline sample)
 *   or_vc_unweighted   jaccard.cpp:236-261 (work = (float)deg), :314-400
 *                      (weight_s = work[u]+work[v]; weight_i += 1.0f per hit),
 *                      :607-620 (out = Wi / (Ws - Wi)), all in fp32 as the
 *                      reference computes them
 *   or_vertex_weights  main.cpp:211-275 (W(k) = sum of row k's edge weights)
 *                      accumulated in fp64
 *   or_weighted_f64    jaccard.cpp:236-261 + :314-400 + :607-620 with vertex
 *                      weights v: S(u) = sum_{k in N(u)} v[k],
 *                      I = sum_{k in N(u) & N(w)} v[k], J = I / (S(u)+S(w)-I),
 *                      evaluated in fp64 (the FP64 host oracle)
 *   or_pairs_*         jaccard.cpp:460-540 semantics per (first[i], second[i])
 *                      pair (cuGraph jaccard_list: the same formula per pair)
 *
 * Parity pins for each function are in tests/test_oracle.py: closed-form
 * graphs (K_n, stars, paths, cycles, K_{a,b}, two triangles sharing an edge),
 * scipy sparse brute force (A@A masked by A; A diag(W) A), networkx
 * jaccard_coefficient, and the committed golden fixtures in tests/golden/.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

int or_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Count |N(a) & N(b)| exactly as jaccard.cpp:715-739: the shorter row is the
 * reference row; each of its columns is binary searched in the longer row.
 * Ni is the degree of `col`'s row in the reference's naming (row = csrInd[tid],
 * col = dest_ind[tid]), so ties pick the edge's source row as the reference. */
static float ec_one(const int32_t *off, const int32_t *ind, int32_t src, int32_t dst) {
  int32_t row = dst, col = src; /* reference: row = csrInd[tid], col = dest_ind[tid] */
  int32_t Ni = off[row + 1] - off[row];
  int32_t Nj = off[col + 1] - off[col];
  int32_t ref = (Ni < Nj) ? row : col;
  int32_t cur = (Ni < Nj) ? col : row;
  float w = 0.0f;
  for (int32_t i = off[ref]; i < off[ref + 1]; i++) {
    int32_t ref_col = ind[i];
    int32_t left = off[cur], right = off[cur + 1] - 1;
    while (left <= right) {
      int32_t middle = (left + right) >> 1;
      int32_t cur_col = ind[middle];
      if (cur_col > ref_col) {
        right = middle - 1;
      } else if (cur_col < ref_col) {
        left = middle + 1;
      } else {
        w = w + 1;
        break;
      }
    }
  }
  return w / ((float)(Ni + Nj) - w);
}

void or_ec_unweighted(int32_t n, int64_t m, const int32_t *off, const int32_t *ind, float *out,
                      int threads) {
  (void)m;
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t u = 0; u < n; u++)
    for (int32_t k = off[u]; k < off[u + 1]; k++) out[k] = ec_one(off, ind, u, ind[k]);
}

/* Per-edge loop over an explicit subset of stored edge ids (sorted or not):
 * the source row of edge k is found by binary search in offsets. */
void or_ec_edges(int32_t n, const int32_t *off, const int32_t *ind, int64_t count,
                 const int64_t *edges, float *out, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t t = 0; t < count; t++) {
    int64_t k = edges[t];
    int32_t lo = 0, hi = n; /* largest u with off[u] <= k */
    while (hi - lo > 1) {
      int32_t mid = (lo + hi) >> 1;
      if (off[mid] <= k) lo = mid; else hi = mid;
    }
    out[t] = ec_one(off, ind, lo, ind[k]);
  }
}

/* Intersection size by merge (used where the reference semantics are the
 * set intersection itself, e.g. pairs of arbitrary vertices). */
static int64_t merge_count(const int32_t *a, int32_t na, const int32_t *b, int32_t nb) {
  int32_t i = 0, j = 0;
  int64_t c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) i++;
    else if (a[i] > b[j]) j++;
    else { c++; i++; j++; }
  }
  return c;
}

static double merge_wsum(const int32_t *a, int32_t na, const int32_t *b, int32_t nb,
                         const double *v) {
  int32_t i = 0, j = 0;
  double s = 0.0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) i++;
    else if (a[i] > b[j]) j++;
    else { s += v[a[i]]; i++; j++; }
  }
  return s;
}

void or_vc_unweighted(int32_t n, int64_t m, const int32_t *off, const int32_t *ind, float *out,
                      int threads) {
  (void)m;
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t u = 0; u < n; u++) {
    float work_u = (float)(off[u + 1] - off[u]); /* RowSum<false>: work[row] = length */
    for (int32_t k = off[u]; k < off[u + 1]; k++) {
      int32_t v = ind[k];
      float work_v = (float)(off[v + 1] - off[v]);
      float ws = work_u + work_v; /* weight_s[j] = work[row] + work[col] */
      float wi = 0.0f;            /* Fill: weight_i = 0; Is: += 1.0f per hit */
      int64_t c = merge_count(ind + off[u], off[u + 1] - off[u], ind + off[v], off[v + 1] - off[v]);
      for (int64_t h = 0; h < c; h++) wi += 1.0f;
      out[k] = wi / (ws - wi); /* Jw: Wi / (Ws - Wi) */
    }
  }
}

void or_vertex_weights(int32_t n, const int32_t *off, const float *edge_w, double *W, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t u = 0; u < n; u++) {
    double s = 0.0;
    for (int32_t k = off[u]; k < off[u + 1]; k++) s += (double)edge_w[k];
    W[u] = s;
  }
}

/* FP64 weighted Jaccard for every stored edge with vertex weights v (fp64). */
void or_weighted_f64(int32_t n, int64_t m, const int32_t *off, const int32_t *ind, const double *v,
                     double *out, int threads) {
  (void)m;
  set_threads(threads);
  double *S = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t u = 0; u < n; u++) {
    double s = 0.0;
    for (int32_t k = off[u]; k < off[u + 1]; k++) s += v[ind[k]];
    S[u] = s;
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t u = 0; u < n; u++)
    for (int32_t k = off[u]; k < off[u + 1]; k++) {
      int32_t w = ind[k];
      double I = merge_wsum(ind + off[u], off[u + 1] - off[u], ind + off[w], off[w + 1] - off[w], v);
      out[k] = I / (S[u] + S[w] - I);
    }
  free(S);
}

/* Pairs, unweighted: out[i] = I / ((float)(d_a + d_b) - I) with I the exact
 * intersection size held in fp32, the per-pair form of the EC arithmetic. */
void or_pairs_unweighted(int32_t n, const int32_t *off, const int32_t *ind, int64_t np,
                         const int32_t *first, const int32_t *second, float *out, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < np; i++) {
    int32_t a = first[i], b = second[i];
    if (a < 0 || a >= n || b < 0 || b >= n) { /* ids outside the graph: NaN */
      out[i] = NAN;
      continue;
    }
    int32_t da = off[a + 1] - off[a], db = off[b + 1] - off[b];
    float I = (float)merge_count(ind + off[a], da, ind + off[b], db);
    out[i] = I / ((float)(da + db) - I);
  }
}

/* Pairs, weighted, fp64. */
void or_pairs_weighted_f64(int32_t n, const int32_t *off, const int32_t *ind, const double *v,
                           int64_t np, const int32_t *first, const int32_t *second, double *out,
                           int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < np; i++) {
    int32_t a = first[i], b = second[i];
    if (a < 0 || a >= n || b < 0 || b >= n) { /* ids outside the graph: NaN */
      out[i] = NAN;
      continue;
    }
    double Sa = 0.0, Sb = 0.0;
    for (int32_t k = off[a]; k < off[a + 1]; k++) Sa += v[ind[k]];
    for (int32_t k = off[b]; k < off[b + 1]; k++) Sb += v[ind[k]];
    double I = merge_wsum(ind + off[a], off[a + 1] - off[a], ind + off[b], off[b + 1] - off[b], v);
    out[i] = I / (Sa + Sb - I);
  }
}

/* Number of outputs above the reference's 1e-5 threshold (jaccard.cpp:811-819). */
int64_t or_nonzero_count(int64_t m, const float *out, float thresh) {
  int64_t c = 0;
  for (int64_t i = 0; i < m; i++)
    if (out[i] > thresh) c++;
  return c;
}
