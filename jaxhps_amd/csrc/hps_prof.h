// Optional per-kernel timing: when enabled (hps_prof_enable), every library
// kernel launch is bracketed by HIP events on its own stream and attributed
// to a kernel class with its algorithmic flop and byte counts.  Disabled by
// default; the launch paths then pay one branch.
#pragma once

#include <hip/hip_runtime.h>

namespace hps {
namespace prof {

bool enabled();

class Scope {
 public:
  Scope(hipStream_t s, const char* name, double flops, double bytes);
  ~Scope();
  Scope(const Scope&) = delete;
  Scope& operator=(const Scope&) = delete;

 private:
  hipStream_t s_;
  const char* name_;
  double flops_, bytes_;
  hipEvent_t e0_ = nullptr;
};

}  // namespace prof
}  // namespace hps
