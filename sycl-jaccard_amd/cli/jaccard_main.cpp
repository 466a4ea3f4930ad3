// jaccard_amd: the command-line driver, counterpart of the reference's
// main.cpp (main.cpp:63-361). Usage and environment are the reference's:
//
//   jaccard_amd <in.mtx|in.csr> <out.mtx|out.csr> [device_index]
//
//   JACCARD_FORCE_EDGE_CENTRIC    edge-centric kernel (main.cpp:48-61)
//   JACCARD_FORCE_VERTEX_CENTRIC  vertex-centric pipeline; wins if both set
//   JACCARD_FORCE_WEIGHTED=0|1    override the file's weightedness
//                                 (main.cpp:139-155); 1 on an unweighted
//                                 graph gives every edge weight 1.0
//   JACCARD_IN_CSR_DUMP_FILEPATH  dump the prepared input CSR (main.cpp:122-135)
//   JACCARD_OUT_CSR_DUMP_FILEPATH dump the result CSR (main.cpp:313-332)
//   JACCARD_EVENT_PROFILE         per-kernel device times on stderr, the
//                                 counterpart of the EVENT_PROFILE build
//
// As in the reference, the edge-centric path computes the unweighted
// coefficients (main.cpp:206) and prints "vertices <n>edges <m>non zero pairs
// <c>" (jaccard.cpp:811-819); the vertex-centric path on a weighted graph
// derives vertex weights from the edge weights first (main.cpp:211-275).
// With neither FORCE variable the reference runs no kernel at all and writes
// uninitialised results (main.cpp:204-209, "TODO: Automatic selection based
// on graph properties" at :203); here jc_select_config picks the algorithm
// and launch tuning from the graph's degree statistics, and the quantity
// follows the input: weighted Jaccard (vertex weights from the edge weights)
// for a weighted graph, unweighted otherwise. Both algorithms give the same
// bits for the same quantity.
//
// Exit codes: 1 bad input extension, 2 bad output extension, 3 binary CSR of
// other types, MTX errors as read_mtx, 4 device or library failure.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/jaccard_amd.h"
#include "../host/graphio.hpp"

namespace {

[[noreturn]] void die(int code, const std::string& msg) {
  std::cerr << msg << std::endl;
  std::exit(code);
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) die(4, std::string(what) + ": " + hipGetErrorString(e));
}

void check_jc(int rc, const char* what) {
  if (rc != JC_OK) die(4, std::string(what) + ": " + jc_strerror(rc));
}

template <typename T>
T* to_device(const std::vector<T>& v) {
  T* d = nullptr;
  check_hip(hipMalloc(&d, std::max<size_t>(v.size(), 1) * sizeof(T)), "hipMalloc");
  if (!v.empty()) check_hip(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return d;
}

// Rows of a CSR read from a file must be sorted and duplicate free for the
// kernels; a file that is not gets its rows sorted (self loops are kept, as
// the reference keeps them for CSR inputs).
void normalize_rows(jcio::Csr& g) {
  bool sorted = true;
  for (int64_t r = 0; r < g.n && sorted; ++r)
    for (int32_t k = g.off[r] + 1; k < g.off[r + 1]; ++k)
      if (g.ind[k - 1] >= g.ind[k]) {
        sorted = false;
        break;
      }
  if (sorted) return;
  std::cerr << "Warning: CSR rows are not sorted and duplicate free; sorting them" << std::endl;
  jcio::Csr h = g;
  h.ind.clear();
  h.w.clear();
  h.off.assign((size_t)g.n + 1, 0);
  std::vector<std::pair<int32_t, float>> row;
  for (int64_t r = 0; r < g.n; ++r) {
    row.clear();
    for (int32_t k = g.off[r]; k < g.off[r + 1]; ++k) row.push_back({g.ind[k], g.w.empty() ? 1.0f : g.w[k]});
    std::stable_sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
    int32_t prev = -1;
    for (auto& e : row) {
      if (e.first == prev) continue;
      prev = e.first;
      h.ind.push_back(e.first);
      if (!g.w.empty()) h.w.push_back(e.second);
    }
    h.off[(size_t)r + 1] = (int32_t)h.ind.size();
  }
  g = std::move(h);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) die(1, "usage: jaccard_amd <in.mtx|in.csr> <out.mtx|out.csr> [device_index]");
  const std::string in_path = argv[1], out_path = argv[2];
  const int in_kind = jcio::file_kind(in_path), out_kind = jcio::file_kind(out_path);
  if (in_kind < 0) die(1, "Input File \"" + in_path + "\" has illegal extension, must be \".mtx\" (text) or \".csr\" (binary)");
  if (out_kind < 0) die(2, "Output File \"" + out_path + "\" has illegal extension, must be \".mtx\" (text) or \".csr\" (binary)");

  // ---- read and prepare the graph (main.cpp:71-120)
  bool weighted = false, directed = false, has_reverse = false, zero_indexed = false;
  jcio::Csr g;
  std::string err;
  if (in_kind == 0) {
    jcio::Coo c;
    // the reference driver reads MTX weights as 1.0 (readMtxToCSR.hpp:29, main.cpp:78)
    const int st = jcio::read_mtx(in_path, c, true, &err);
    if (st != jcio::kOk) die(st, err);
    weighted = c.weighted;
    directed = c.directed;
    g = jcio::csr_from_coo(c, !directed);
    has_reverse = !directed;
  } else {
    const int st = jcio::read_csr(in_path, g, &err);
    if (st == jcio::kErrTypes) die(3, "Binary CSR Input Header does not match required data types");
    if (st != jcio::kOk) die(st, err);
    weighted = g.flags & jcio::kWeighted;
    directed = g.flags & jcio::kDirected;
    has_reverse = g.flags & jcio::kHasReverseEdges;
    zero_indexed = g.flags & jcio::kZeroIndexed;
    if (!directed && !has_reverse) {
      jcio::Coo c = jcio::coo_from_csr(g);
      c.weighted = weighted;
      g = jcio::csr_from_coo(c, true);
      has_reverse = true;
    } else {
      normalize_rows(g);
    }
  }
  if (const char* p = std::getenv("JACCARD_IN_CSR_DUMP_FILEPATH")) {
    jcio::Csr d = g;
    d.flags = (weighted ? (uint64_t)jcio::kWeighted : 0) | jcio::kHasReverseEdges;
    if (!weighted) d.w.clear();
    if (jcio::write_csr(p, d, &err) != jcio::kOk) std::cerr << err << std::endl;
  }
  if (const char* p = std::getenv("JACCARD_FORCE_WEIGHTED")) {
    const bool had = weighted;
    if (std::strcmp(p, "1") == 0) weighted = true;
    if (std::strcmp(p, "0") == 0) weighted = false;
    if (weighted && !had) g.w.assign(g.ind.size(), 1.0f);
  }
  if (weighted && g.w.size() != g.ind.size()) g.w.assign(g.ind.size(), 1.0f);
  const int32_t n = (int32_t)g.n, m = (int32_t)g.ind.size();

  // ---- device (main.cpp:157-187)
  int ndev = 0;
  check_hip(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  int dev = 0;
  if (argc >= 4) {
    for (int i = 0; i < ndev; ++i) {
      hipDeviceProp_t pr;
      check_hip(hipGetDeviceProperties(&pr, i), "hipGetDeviceProperties");
      std::cerr << "HIP Device [" << i << "]: " << pr.name << std::endl;
    }
    dev = std::atoi(argv[3]);
    if (dev < 0 || dev >= ndev) die(4, "device index out of range");
  }
  check_hip(hipSetDevice(dev), "hipSetDevice");
  hipDeviceProp_t prop;
  check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  std::cout << "Running on " << prop.name << "\n";

  // ---- implementation choice (main.cpp:48-61, 202-209)
  bool ec = false, vc = false;
  if (std::getenv("JACCARD_FORCE_EDGE_CENTRIC")) {
    std::cerr << "FORCE Edge-Centric Implementation" << std::endl;
    ec = true;
  }
  if (std::getenv("JACCARD_FORCE_VERTEX_CENTRIC")) {
    std::cerr << "FORCE Vertex-Centric Implementation" << std::endl;
    vc = true;
    ec = false;
  }
  const bool autosel = !ec && !vc;
  const bool profile = std::getenv("JACCARD_EVENT_PROFILE") != nullptr;
  if (profile) jc_profile_enable(1);

  int32_t* d_off = to_device(g.off);
  int32_t* d_ind = to_device(g.ind);
  float* d_ew = weighted ? to_device(g.w) : nullptr;
  float* d_out = nullptr;
  check_hip(hipMalloc(&d_out, std::max<size_t>((size_t)m, 1) * sizeof(float)), "hipMalloc");
  jc_csr c{n, m, d_off, d_ind, d_ew};
  jc_tuning tuning{};
  if (autosel) {
    int e = JC_EDGE_CENTRIC;
    check_jc(jc_select_config(&c, weighted ? 1 : 0, &e, &tuning, nullptr), "jc_select_config");
    ec = e == JC_EDGE_CENTRIC;
    vc = !ec;
    std::cerr << "AUTO " << (ec ? "Edge-Centric" : "Vertex-Centric") << " Implementation"
              << (weighted ? " (weighted)" : "") << " small_max " << tuning.small_max << " small_blocks_per_cu "
              << tuning.small_blocks_per_cu << " bitmap_blocks_per_cu " << tuning.bitmap_blocks_per_cu
              << " hash_table " << tuning.hash_table << std::endl;
  }
  // vertex weights: the weighted quantity of the vertex-centric path, and of
  // the automatic choice on a weighted graph (a forced edge-centric run is
  // unweighted, as in the reference)
  float* d_w = nullptr;
  if (weighted && (vc || autosel)) {
    check_hip(hipMalloc(&d_w, std::max<size_t>((size_t)n, 1) * sizeof(float)), "hipMalloc");
    check_jc(jc_vertex_weights(&c, d_w, nullptr), "jc_vertex_weights");
  }
  const jc_tuning* tp = autosel ? &tuning : nullptr;
  if (ec) {
    check_jc(jc_jaccard_tuned(&c, d_w, d_out, JC_EDGE_CENTRIC, 0u, nullptr, 0, tp, nullptr), "jc_jaccard");
    int64_t nz = 0;
    check_jc(jc_nonzero_count(m, d_out, 1e-5f, &nz, nullptr), "jc_nonzero_count");
    std::cout << "vertices " << n << "edges " << m << "non zero pairs " << nz << std::endl;
  } else {
    check_jc(jc_jaccard_tuned(&c, d_w, d_out, JC_VERTEX_CENTRIC, 0u, nullptr, 0, tp, nullptr), "jc_jaccard");
  }
  if (d_w) check_hip(hipFree(d_w), "hipFree");
  std::vector<float> res((size_t)m);
  if (m) check_hip(hipMemcpy(res.data(), d_out, (size_t)m * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
  check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
  if (profile) {
    char names[512];
    jc_profile_names(names, sizeof(names));
    for (char* tok = std::strtok(names, ","); tok; tok = std::strtok(nullptr, ",")) {
      int cnt = 0;
      double ms = 0.0;
      if (jc_profile_query(tok, &cnt, &ms) == JC_OK && cnt > 0)
        std::cerr << tok << " kernel elapsed time: " << (long long)(ms * 1e6) << " (ns)" << std::endl;
    }
  }
  (void)hipFree(d_off);
  (void)hipFree(d_ind);
  if (d_ew) (void)hipFree(d_ew);
  (void)hipFree(d_out);

  // ---- results as edge data (main.cpp:277-358)
  jcio::Csr r = g;
  r.w = res;
  r.flags = jcio::kWeighted | (zero_indexed ? (uint64_t)jcio::kZeroIndexed : 0) |
            (directed ? (uint64_t)jcio::kDirected : 0) | (has_reverse ? (uint64_t)jcio::kHasReverseEdges : 0);
  jcio::Coo rc;
  bool reverse_removed = false;
  if (has_reverse && out_kind == 0 && !directed) {
    rc = jcio::coo_from_csr(r);
    jcio::remove_reverse(rc);
    reverse_removed = true;
  }
  if (const char* p = std::getenv("JACCARD_OUT_CSR_DUMP_FILEPATH")) {
    jcio::Csr d = r;
    if (reverse_removed) {
      jcio::Coo t = rc;
      t.weighted = true;
      d = jcio::csr_from_coo(t, false);
      d.flags = jcio::kWeighted | (directed ? (uint64_t)jcio::kDirected : 0);
    }
    if (jcio::write_csr(p, d, &err) != jcio::kOk) std::cerr << err << std::endl;
  }
  int st;
  if (out_kind == 0) {
    if (!reverse_removed) rc = jcio::coo_from_csr(r);
    rc.n = n;
    st = jcio::write_mtx(out_path, rc, true, directed, &err);
  } else {
    st = jcio::write_csr(out_path, r, &err);
  }
  if (st != jcio::kOk) die(4, err);
  return 0;
}
