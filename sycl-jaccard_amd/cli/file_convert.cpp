// file_convert: converts graphs between the MatrixMarket and binary CSR
// formats, counterpart of the reference's fileConvert.cpp (fileConvert.cpp:27-174).
//
//   file_convert <in.mtx|in.csr> <out.mtx|out.csr> <keepReverseEdges 0|1>
//
//   CONVERT_FORCE_DROP_WEIGHTS  read MTX weights as absent
//   CONVERT_FORCE_REVERSE       generate reverse edges even where the rules
//                               below would not keep them
//
// keepReverseEdges asks for both directions of every undirected edge in the
// output. It is turned off for directed inputs and for MTX outputs (where a
// symmetric file lists one direction), unless CONVERT_FORCE_REVERSE is set.
#include <cstdlib>
#include <iostream>
#include <string>

#include "../host/graphio.hpp"

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cerr << "Error, incorrect number of args, usage is:\n./file_convert <input.[mtx|csr]> "
                 "<output.[mtx|csr]> <keepReverseEdges (0 or 1)>"
              << std::endl;
    return 1;
  }
  const std::string in_path = argv[1], out_path = argv[2];
  const int in_kind = jcio::file_kind(in_path), out_kind = jcio::file_kind(out_path);
  if (in_kind < 0) {
    std::cerr << "Input File \"" << in_path << "\" has illegal extension, must be \".mtx\" (text) or \".csr\" (binary)"
              << std::endl;
    return 1;
  }
  if (out_kind < 0) {
    std::cerr << "Output File \"" << out_path
              << "\" has illegal extension, must be \".mtx\" (text) or \".csr\" (binary)" << std::endl;
    return 2;
  }
  bool keep_reverse = std::atoi(argv[3]) != 0;
  const bool drop_weights = std::getenv("CONVERT_FORCE_DROP_WEIGHTS") != nullptr;
  if (drop_weights) std::cerr << "FORCE Drop Weights" << std::endl;
  std::string err;
  jcio::Coo coo;
  bool weighted = false, directed = false, has_reverse = false;
  if (in_kind == 0) {
    const int st = jcio::read_mtx(in_path, coo, drop_weights, &err);
    if (st != jcio::kOk) {
      std::cerr << err << std::endl;
      return st;
    }
    weighted = coo.weighted && !drop_weights;
    directed = coo.directed;
  } else {
    jcio::Csr g;
    const int st = jcio::read_csr(in_path, g, &err);
    if (st != jcio::kOk) {
      std::cerr << err << std::endl;
      return st;
    }
    coo = jcio::coo_from_csr(g);
    weighted = g.flags & jcio::kWeighted;
    directed = g.flags & jcio::kDirected;
    has_reverse = g.flags & jcio::kHasReverseEdges;
  }
  coo.weighted = weighted;
  coo.directed = directed;
  if (keep_reverse) {
    if (directed) {
      std::cerr << "Warning, Cannot retain reverseEdges of Directed input, could cause collisions" << std::endl;
      keep_reverse = false;
    }
    if (out_kind == 0) {
      std::cerr << "Warning, Cannot retain reverseEdges with MTX output, would be indistinguishable from directed"
                << std::endl;
      keep_reverse = false;
    }
    if (std::getenv("CONVERT_FORCE_REVERSE")) {
      std::cerr << "FORCE Reverse Edge Generation" << std::endl;
      keep_reverse = true;
    }
  }
  jcio::Csr out;
  if (keep_reverse && !has_reverse) {
    out = jcio::csr_from_coo(coo, true);
    has_reverse = true;
  } else {
    if (has_reverse && !keep_reverse) {
      jcio::remove_reverse(coo);
      has_reverse = false;
    }
    out = jcio::csr_from_coo(coo, false);
  }
  out.flags = (weighted ? (uint64_t)jcio::kWeighted : 0) | jcio::kZeroIndexed |
              (directed ? (uint64_t)jcio::kDirected : 0) | (has_reverse ? (uint64_t)jcio::kHasReverseEdges : 0);
  if (!weighted) out.w.clear();
  int st;
  if (out_kind == 1) {
    st = jcio::write_csr(out_path, out, &err);
  } else {
    jcio::Coo c = jcio::coo_from_csr(out);
    st = jcio::write_mtx(out_path, c, weighted, directed, &err);
  }
  if (st != jcio::kOk) std::cerr << err << std::endl;
  return st;
}
