// Per-kernel device timing, the counterpart of the reference's EVENT_PROFILE
// build (standalone_algorithms.hpp:37-47 wait_and_print; used at
// jaccard.cpp:798-809, 875-896 and main.cpp:236-238, 270-272), which prints
// "<Kernel> kernel elapsed time: <ns> (ns)" per kernel.
//
// Here it is a runtime switch: when enabled, every kernel launch of the
// library is bracketed by two hipEvents recorded on the launch stream; the
// query sums the elapsed times per kernel name. Disabled, a scope costs one
// branch.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "jc_profile.hpp"

namespace jc {
namespace {

struct Rec {
  std::string name;
  hipEvent_t a, b;
};

std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

}  // namespace

bool profiling_on() { return g_on; }

ProfScope::ProfScope(const char* name, hipStream_t s) : name_(name), s_(s) {
  if (!g_on) return;
  std::lock_guard<std::mutex> lk(g_mu);
  a_ = take_event();
  if (a_) (void)hipEventRecord(a_, s_);
}

ProfScope::~ProfScope() {
  if (!g_on || !a_) return;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t b = take_event();
  if (!b) return;
  (void)hipEventRecord(b, s_);
  g_recs.push_back(Rec{name_, a_, b});
}

}  // namespace jc

extern "C" {

int jc_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(jc::g_mu);
  jc::g_on = on != 0;
  return 0;
}

int jc_profile_reset(void) {
  std::lock_guard<std::mutex> lk(jc::g_mu);
  for (auto& r : jc::g_recs) {
    (void)hipEventSynchronize(r.b);
    jc::g_pool.push_back(r.a);
    jc::g_pool.push_back(r.b);
  }
  jc::g_recs.clear();
  return 0;
}

// Synchronous: waits for the recorded events. Writes launch count and total
// milliseconds of kernel `name` (all kernels if name is NULL or empty).
int jc_profile_query(const char* name, int* count, double* total_ms) {
  std::lock_guard<std::mutex> lk(jc::g_mu);
  int c = 0;
  double t = 0.0;
  for (auto& r : jc::g_recs) {
    if (name && *name && r.name != name) continue;
    if (hipEventSynchronize(r.b) != hipSuccess) return -4;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return -4;
    ++c;
    t += ms;
  }
  if (count) *count = c;
  if (total_ms) *total_ms = t;
  return 0;
}

// Comma-separated distinct kernel names recorded so far (truncated to len).
int jc_profile_names(char* buf, size_t len) {
  std::lock_guard<std::mutex> lk(jc::g_mu);
  std::map<std::string, int> seen;
  std::string s;
  for (auto& r : jc::g_recs)
    if (seen.emplace(r.name, 1).second) s += (s.empty() ? "" : ",") + r.name;
  if (buf && len) {
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return (int)s.size();
}

}  // extern "C"
