// Per-kernel event timing scope (see jc_profile.cpp).
#pragma once
#include <hip/hip_runtime.h>

namespace jc {

bool profiling_on();

class ProfScope {
 public:
  ProfScope(const char* name, hipStream_t s);
  ~ProfScope();
  ProfScope(const ProfScope&) = delete;
  ProfScope& operator=(const ProfScope&) = delete;

 private:
  const char* name_;
  hipStream_t s_;
  hipEvent_t a_ = nullptr;
};

}  // namespace jc
