// Graph file formats of the reference workflow and the host-side graph
// preparation around the Jaccard call.
//
//   * MatrixMarket coordinate text (.mtx), 1-based, "general" = directed,
//     anything else = undirected (readMtxToCSR.cpp:48-109).
//   * The binary CSR format, version 2 (.csr): a 32-byte header {int64
//     version, int64 vertices, int64 edges, 64 flag bits}, then offsets,
//     indices and, if weighted, edge weights (readMtxToCSR.hpp:44-64,
//     readMtxToCSR.cpp:285-395). Flag bits from bit 0: weighted,
//     zero-indexed, directed, has-reverse-edges, 64-bit vertex ids, 64-bit
//     edge offsets, 64-bit weights.
//
// Deliberate differences from the reference (its quirks are not reproduced):
// the vertex count is max(header size, largest id + 1) and every vertex is
// kept (the reference drops trailing isolated vertices and shifts rows when
// the lowest id is isolated, readMtxToCSR.cpp:152-179); a repeated (src, dst)
// keeps its first weight (the reference keeps one CSR entry per distinct
// (src, dst, weight)).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace jcio {

constexpr int64_t kCsrVersion = 2;

enum CsrFlag : uint64_t {
  kWeighted = 1u << 0,
  kZeroIndexed = 1u << 1,
  kDirected = 1u << 2,
  kHasReverseEdges = 1u << 3,
  kVertexT64 = 1u << 4,
  kEdgeT64 = 1u << 5,
  kWeightT64 = 1u << 6,
};

struct CsrHeader {
  int64_t version;
  int64_t num_verts;
  int64_t num_edges;
  uint64_t flags;
};
static_assert(sizeof(CsrHeader) == 32, "binary CSR header is 32 bytes");

// Coordinate list, 0-based ids.
struct Coo {
  int64_t n = 0;
  std::vector<int32_t> src, dst;
  std::vector<float> w;  // one per entry (1.0 when the file has none)
  bool weighted = false;
  bool directed = false;
};

// In-memory CSR with the file's facts.
struct Csr {
  int64_t n = 0;
  std::vector<int32_t> off, ind;
  std::vector<float> w;  // empty when unweighted
  uint64_t flags = 0;
};

// Error codes (also the exit codes of the command line tools, as in the
// reference: bad extension 1/2, non-matrix / non-coordinate / complex MTX
// 1/2/3, CSR type mismatch 3).
enum Status : int {
  kOk = 0,
  kErrOpen = 10,
  kErrNotMatrix = 1,
  kErrNotCoordinate = 2,
  kErrComplex = 3,
  kErrTypes = 3,
  kErrFormat = 11,
};

// .mtx / .csr by extension (0 = mtx, 1 = csr, -1 = other).
int file_kind(const std::string& path);

int read_mtx(const std::string& path, Coo& out, bool drop_weights, std::string* err);
int write_mtx(const std::string& path, const Coo& coo, bool weighted, bool directed, std::string* err);

int read_csr(const std::string& path, Csr& out, std::string* err);
int write_csr(const std::string& path, const Csr& g, std::string* err);

// Builds a sorted, deduplicated CSR without self loops; add_reverse inserts
// (v, u, w) for every (u, v, w).
Csr csr_from_coo(const Coo& coo, bool add_reverse);
// Every stored entry as a coordinate (row-major order).
Coo coo_from_csr(const Csr& g);
// Keeps one direction of each undirected pair: the entry with src > dst
// (and self loops), the lower triangle of a symmetric MatrixMarket file.
void remove_reverse(Coo& coo);

}  // namespace jcio
