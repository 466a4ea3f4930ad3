// sygraph_jaccard.hpp -- C++ drop-in for the reference's algorithm API.
//
// Restores the call shapes of standalone_algorithms.hpp:67-69, :86-87,
// :109-112, :131-133 (sygraph::jaccard / sygraph::jaccard_list) and the
// GraphCSRView container of standalone_csr.hpp:44-105 on top of the C ABI in
// jaccard_amd.h. Differences, all forced by HIP instead of SYCL:
//   * buffers are device pointers (cl::sycl::buffer<T> is replaced by T*);
//   * the queue is a hipStream_t (cl::sycl::queue& is replaced by a stream);
//   * failures throw sygraph::jaccard_error (the reference prints SYCL
//     exceptions and carries on, jaccard.cpp:178-180).
// Only <int32_t, int32_t, float>, the instantiation the reference builds
// (Makefile:33, jaccard.cpp:990-1021), is provided.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "jaccard_amd.h"

namespace sygraph {

class jaccard_error : public std::runtime_error {
 public:
  jaccard_error(int code, const std::string& what)
      : std::runtime_error(what + ": " + jc_strerror(code)), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

// Non-owning view of a device-resident CSR graph (GraphCSRView,
// standalone_csr.hpp:44-105): offsets[n+1], indices[m], optional edge_data[m].
template <typename vertex_t, typename edge_t, typename weight_t>
struct GraphCSRView {
  using vertex_type = vertex_t;
  using edge_type = edge_t;
  using weight_type = weight_t;
  weight_t* edge_data = nullptr;
  vertex_t number_of_vertices = 0;
  edge_t number_of_edges = 0;
  edge_t* offsets = nullptr;
  vertex_t* indices = nullptr;

  GraphCSRView() = default;
  GraphCSRView(edge_t* offsets_, vertex_t* indices_, weight_t* edge_data_, vertex_t n, edge_t m)
      : edge_data(edge_data_), number_of_vertices(n), number_of_edges(m), offsets(offsets_), indices(indices_) {}
};

namespace detail {
template <typename VT, typename ET, typename WT>
inline jc_csr to_c(const GraphCSRView<VT, ET, WT>& g) {
  static_assert(std::is_same<VT, int32_t>::value && std::is_same<ET, int32_t>::value &&
                    std::is_same<WT, float>::value,
                "only GraphCSRView<int32_t, int32_t, float> is built, as in the reference");
  return jc_csr{g.number_of_vertices, g.number_of_edges, g.offsets, g.indices, g.edge_data};
}
inline void check(int code, const char* what) {
  if (code != JC_OK) throw jaccard_error(code, what);
}
}  // namespace detail

// Weighted (vertex weights, n entries) Jaccard of every stored edge
// (standalone_algorithms.hpp:67-69). result: m entries, caller-allocated.
template <bool edge_centric, typename VT, typename ET, typename WT>
void jaccard(GraphCSRView<VT, ET, WT>& graph, WT* weights, WT* result, jc_stream_t stream = nullptr) {
  const jc_csr g = detail::to_c(graph);
  detail::check(jc_jaccard(&g, weights, result, edge_centric ? JC_EDGE_CENTRIC : JC_VERTEX_CENTRIC, stream),
                "sygraph::jaccard");
}

// Unweighted Jaccard of every stored edge (standalone_algorithms.hpp:86-87).
template <bool edge_centric, typename VT, typename ET, typename WT>
void jaccard(GraphCSRView<VT, ET, WT>& graph, WT* result, jc_stream_t stream = nullptr) {
  const jc_csr g = detail::to_c(graph);
  detail::check(jc_jaccard(&g, nullptr, result, edge_centric ? JC_EDGE_CENTRIC : JC_VERTEX_CENTRIC, stream),
                "sygraph::jaccard");
}

// Weighted Jaccard of (first[i], second[i]) pairs (standalone_algorithms.hpp:109-112).
template <typename VT, typename ET, typename WT>
void jaccard_list(GraphCSRView<VT, ET, WT>& graph, WT* weights, ET num_pairs, VT* first, VT* second, WT* result,
                  jc_stream_t stream = nullptr) {
  const jc_csr g = detail::to_c(graph);
  detail::check(jc_jaccard_list(&g, weights, num_pairs, first, second, result, stream), "sygraph::jaccard_list");
}

// Unweighted Jaccard of (first[i], second[i]) pairs (standalone_algorithms.hpp:131-133).
template <typename VT, typename ET, typename WT>
void jaccard_list(GraphCSRView<VT, ET, WT>& graph, ET num_pairs, VT* first, VT* second, WT* result,
                  jc_stream_t stream = nullptr) {
  const jc_csr g = detail::to_c(graph);
  detail::check(jc_jaccard_list(&g, nullptr, num_pairs, first, second, result, stream), "sygraph::jaccard_list");
}

// The driver's weighted preamble (main.cpp:211-275): W[k] = sum of row k's
// edge weights, from graph.edge_data.
template <typename VT, typename ET, typename WT>
void vertex_weights(GraphCSRView<VT, ET, WT>& graph, WT* W, jc_stream_t stream = nullptr) {
  const jc_csr g = detail::to_c(graph);
  detail::check(jc_vertex_weights(&g, W, stream), "sygraph::vertex_weights");
}

}  // namespace sygraph
