// Host program written against the C++ drop-in header (sygraph_jaccard.hpp),
// used by the tests to show that the header compiles and that every call
// shape of the reference's algorithm API (standalone_algorithms.hpp:67-69,
// :86-87, :109-112, :131-133, and the driver's vertex-weight preamble,
// main.cpp:211-275) reaches the library.
//
//   dropin_check --errors            argument errors only (no device use):
//                                    each call must throw sygraph::jaccard_error
//                                    with JC_ERR_INVALID_ARG; exit 0 if so
//   dropin_check <in.bin> <out.bin>  runs every overload on the graph in
//                                    in.bin and writes the results to out.bin
//
// in.bin (little endian): int32 n, m, k; int32 offsets[n+1], indices[m];
// float edge_w[m], vertex_w[n]; int32 first[k], second[k].
// out.bin: float W[n] (vertex_weights), then m floats each for
// jaccard<true>(weights), jaccard<false>(weights), jaccard<true>(),
// jaccard<false>(), then k floats each for jaccard_list(weights) and
// jaccard_list().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sygraph_jaccard.hpp"

using View = sygraph::GraphCSRView<int32_t, int32_t, float>;

static int fail(const char* what) {
  std::fprintf(stderr, "dropin_check: %s\n", what);
  return 1;
}

// Every overload with an invalid graph (negative vertex count): the library
// rejects it before touching the device.
static int check_errors() {
  View g(nullptr, nullptr, nullptr, -1, 0);
  float* none = nullptr;
  int32_t* ids = nullptr;
  int thrown = 0;
  auto expect = [&](auto&& call) {
    try {
      call();
    } catch (const sygraph::jaccard_error& e) {
      if (e.code() == JC_ERR_INVALID_ARG) ++thrown;
    }
  };
  expect([&] { sygraph::jaccard<true>(g, none, none); });
  expect([&] { sygraph::jaccard<false>(g, none, none); });
  expect([&] { sygraph::jaccard<true>(g, none); });
  expect([&] { sygraph::jaccard<false>(g, none); });
  expect([&] { sygraph::jaccard_list(g, none, 1, ids, ids, none); });
  expect([&] { sygraph::jaccard_list(g, 1, ids, ids, none); });
  expect([&] { sygraph::vertex_weights(g, none); });
  if (thrown != 7) {
    std::fprintf(stderr, "dropin_check: %d of 7 calls threw JC_ERR_INVALID_ARG\n", thrown);
    return 1;
  }
  std::printf("errors ok: 7 of 7 calls threw sygraph::jaccard_error(JC_ERR_INVALID_ARG)\n");
  return 0;
}

template <typename T>
static bool read_n(FILE* f, std::vector<T>& v, size_t n) {
  v.resize(n);
  return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

template <typename T>
static T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  if (hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(T)) != hipSuccess) return nullptr;
  if (!h.empty() && hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  return d;
}

int main(int argc, char** argv) {
  if (argc == 2 && std::string(argv[1]) == "--errors") return check_errors();
  if (argc != 3) return fail("usage: dropin_check --errors | dropin_check <in.bin> <out.bin>");
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return fail("cannot open input");
  int32_t hdr[3];
  if (std::fread(hdr, sizeof(int32_t), 3, f) != 3) return fail("short header");
  const int32_t n = hdr[0], m = hdr[1], k = hdr[2];
  std::vector<int32_t> off, ind, first, second;
  std::vector<float> ew, vw;
  if (!read_n(f, off, (size_t)n + 1) || !read_n(f, ind, m) || !read_n(f, ew, m) || !read_n(f, vw, n) ||
      !read_n(f, first, k) || !read_n(f, second, k))
    return fail("short input");
  std::fclose(f);

  int32_t* d_off = upload(off);
  int32_t* d_ind = upload(ind);
  float* d_ew = upload(ew);
  float* d_vw = upload(vw);
  int32_t* d_first = upload(first);
  int32_t* d_second = upload(second);
  float *d_W = nullptr, *d_out = nullptr;
  if (!d_off || !d_ind || !d_ew || !d_vw || !d_first || !d_second ||
      hipMalloc(&d_W, std::max(n, 1) * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_out, std::max(std::max(m, k), 1) * sizeof(float)) != hipSuccess)
    return fail("device allocation");

  View g(d_off, d_ind, d_ew, n, m);
  std::vector<float> res;
  auto take = [&](size_t count) {
    if (hipDeviceSynchronize() != hipSuccess) return false;
    const size_t at = res.size();
    res.resize(at + count);
    return count == 0 || hipMemcpy(res.data() + at, d_out, count * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
  };
  try {
    sygraph::vertex_weights(g, d_W);
    if (hipDeviceSynchronize() != hipSuccess) return fail("vertex_weights");
    res.resize(n);
    if (n && hipMemcpy(res.data(), d_W, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return fail("copy W");
    sygraph::jaccard<true>(g, d_vw, d_out);
    if (!take(m)) return fail("jaccard<true> weighted");
    sygraph::jaccard<false>(g, d_vw, d_out);
    if (!take(m)) return fail("jaccard<false> weighted");
    sygraph::jaccard<true>(g, d_out);
    if (!take(m)) return fail("jaccard<true>");
    sygraph::jaccard<false>(g, d_out);
    if (!take(m)) return fail("jaccard<false>");
    sygraph::jaccard_list(g, d_vw, k, d_first, d_second, d_out);
    if (!take(k)) return fail("jaccard_list weighted");
    sygraph::jaccard_list(g, k, d_first, d_second, d_out);
    if (!take(k)) return fail("jaccard_list");
  } catch (const sygraph::jaccard_error& e) {
    std::fprintf(stderr, "dropin_check: %s\n", e.what());
    return 2;
  }
  FILE* o = std::fopen(argv[2], "wb");
  if (!o || std::fwrite(res.data(), sizeof(float), res.size(), o) != res.size()) return fail("cannot write output");
  std::fclose(o);
  for (void* p : {(void*)d_off, (void*)d_ind, (void*)d_ew, (void*)d_vw, (void*)d_first, (void*)d_second,
                  (void*)d_W, (void*)d_out})
    (void)hipFree(p);
  std::printf("dropin ok: n=%d m=%d k=%d\n", n, m, k);
  return 0;
}
