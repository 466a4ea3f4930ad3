// compare_coords: compares two Jaccard result files coordinate by coordinate,
// counterpart of the reference's compareCoords.cpp (compareCoords.cpp:174-320).
//
//   compare_coords <gold.mtx|gold.csr> <test.mtx|test.csr> <tolerance>
//
// Both files are read with their values (MTX weights are kept); the test
// file is converted to the gold file's layout. Entries present in one file
// only, and values differing by more than the tolerance, are reported on
// stderr. Prints "Inputs successfully read, proceeding to test" and progress
// on stdout. Exit code: the number of reported differences, capped at 255
// (the reference returns the count itself, which the shell truncates to 8
// bits).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../host/graphio.hpp"

namespace {

struct Entry {
  int64_t s, d;
  float w;
};

// Entries of a file sorted by (src, dst), ids as in the file (MTX 1-based,
// CSR 0-based unless the file says otherwise).
int load(const std::string& path, std::vector<Entry>& out, int64_t& n, int& kind) {
  std::string err;
  kind = jcio::file_kind(path);
  if (kind == 0) {
    jcio::Coo c;
    const int st = jcio::read_mtx(path, c, false, &err);
    if (st != jcio::kOk) {
      std::cerr << err << std::endl;
      return st;
    }
    n = c.n;
    for (size_t k = 0; k < c.src.size(); ++k) out.push_back({(int64_t)c.src[k] + 1, (int64_t)c.dst[k] + 1, c.w[k]});
  } else if (kind == 1) {
    jcio::Csr g;
    const int st = jcio::read_csr(path, g, &err);
    if (st != jcio::kOk) {
      std::cerr << (st == jcio::kErrTypes ? "Binary CSR Input Header does not match required weight type" : err)
                << std::endl;
      return st;
    }
    n = g.n;
    for (int64_t r = 0; r < g.n; ++r)
      for (int32_t k = g.off[r]; k < g.off[r + 1]; ++k) out.push_back({r, g.ind[k], g.w.empty() ? 1.0f : g.w[k]});
  } else {
    std::cerr << "File extension must be csr or mtx" << std::endl;
    return 1;
  }
  std::stable_sort(out.begin(), out.end(), [](const Entry& a, const Entry& b) { return std::tie(a.s, a.d) < std::tie(b.s, b.d); });
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cerr << "Error, incorrect number of args, usage is:\n./compare_coords <goldFile> <testFile> <DP_tolerance>"
              << std::endl;
    return 1;
  }
  std::vector<Entry> gold, test;
  int64_t ng = 0, nt = 0;
  int kg = -1, kt = -1;
  if (int st = load(argv[1], gold, ng, kg)) return st;
  if (int st = load(argv[2], test, nt, kt)) return st;
  if (kg != kt) {  // bring the test file's ids to the gold file's base
    const int64_t shift = kg == 0 ? 1 : -1;
    for (auto& e : test) e.s += shift, e.d += shift;
  }
  const double tol = std::atof(argv[3]);
  std::cout << "Inputs successfully read, proceeding to test" << std::endl;
  const char* base = kg == 0 ? "one" : "zero";
  if (gold.size() != test.size())
    std::cerr << "Warning: sets differ in size! Gold: " << gold.size() << " vs. Test: " << test.size() << std::endl;
  long warn = 0;
  size_t i = 0, j = 0, count = 0;
  const size_t one_pct = std::max<size_t>(gold.size() / 100, 1);
  while (i < gold.size() && j < test.size()) {
    const Entry &a = gold[i], &b = test[j];
    if (a.s == b.s && a.d == b.d) {
      if (std::fabs((double)a.w - (double)b.w) > tol) {
        std::cerr << "Warning: " << base << "-indexed elements at (" << a.s << "," << a.d << ") differ by more than "
                  << tol << " tolerance! Gold: " << a.w << " vs. Test: " << b.w << std::endl;
        ++warn;
      }
      ++i, ++j;
    } else if (std::tie(a.s, a.d) < std::tie(b.s, b.d)) {
      std::cerr << "Element missing from test file: " << a.s << " " << a.d << " " << a.w << std::endl;
      ++warn, ++i;
    } else {
      std::cerr << "Element added to test file: " << b.s << " " << b.d << " " << b.w << std::endl;
      ++warn, ++j;
    }
    if (++count % one_pct == 0) std::cout << "Completed " << count / one_pct << "% of scan" << std::endl;
  }
  if (i != gold.size()) {
    std::cerr << "Warning: " << gold.size() - i << " elements unscanned from Gold input" << std::endl;
    warn += (long)(gold.size() - i);
  }
  if (j != test.size()) {
    std::cerr << "Warning: " << test.size() - j << " elements unscanned from Test input" << std::endl;
    warn += (long)(test.size() - j);
  }
  return (int)std::min<long>(warn, 255);
}
