// Internal to libhps_mi355x: the handle, argument/error macros and launch
// helpers shared by the C-ABI translation units.
#pragma once

#include <string>

#include "hps_kernels.h"
#include "hps_prof.h"
#include "../../include/hps_mi355x.h"

// Open a timing scope named "<label><dtype>" for the next launch (no-op unless
// profiling is enabled).
#define HPS_PROF(label, T, flops, bytes)                                                 \
  static const std::string _pn_##__LINE__ = std::string(label "<") + tname<T>() + ">"; \
  prof::Scope _ps_##__LINE__(s, _pn_##__LINE__.c_str(), (flops), (bytes))

namespace hps {

// coefficient k multiplies operator op[k] of the stacked diff operators; -1 is the identity
struct CoefOps {
  int op[16];
};

inline dim3 grid_for(long total, int batch) {
  unsigned gx = ceil_div(total, 256);
  if (gx > 2048) gx = 2048;
  if (gx == 0) gx = 1;
  unsigned gy = batch < 65535 ? batch : 65535;
  return dim3(gx, gy);
}

}  // namespace hps

struct hps_handle_s {
  hipStream_t stream = nullptr;
  int device = -1;  // the stream's device at creation
  bool merge_nopiv = false;  // hps_merge_level factors without row exchanges
  bool leaf_nopiv = false;   // hps_leaf_solve factors without row exchanges (checked)
  std::string err;
};

using namespace hps;

#define HPS_CHECK_HIP(h, expr)                                                      \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      (h)->err = std::string(#expr) + ": " + hipGetErrorString(_e);                 \
      return HPS_ERR_HIP;                                                           \
    }                                                                               \
  } while (0)

#define HPS_ARG(h, cond, msg)   \
  do {                          \
    if (!(cond)) {              \
      (h)->err = (msg);         \
      return HPS_ERR_ARG;       \
    }                           \
  } while (0)

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

#define DISPATCH(dtype, FN, ...)                                  \
  switch (dtype) {                                                \
    case HPS_F64: return FN<double>(__VA_ARGS__);                 \
    case HPS_F32: return FN<float>(__VA_ARGS__);                  \
    case HPS_C128: return FN<c128>(__VA_ARGS__);                  \
    case HPS_C64: return FN<c64>(__VA_ARGS__);                    \
    default: h->err = "bad dtype"; return HPS_ERR_ARG;            \
  }

