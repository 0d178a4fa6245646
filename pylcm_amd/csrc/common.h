// Shared helpers of the pylcm_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pylcm_amd.h"

struct lcm_ctx {
  hipStream_t stream;
  int device;
};

namespace lcm {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

#define LCM_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return ::lcm::check_hip(_e, #call); \
  } while (0)

// Zero weight drops its term whatever value stands there; otherwise w * v.
// Mirrors zero_safe_weighted_term (src/_lcm/zero_safe.py:161-251) for weights
// in the normal range.
template <typename T>
__device__ __forceinline__ T zero_safe(T w, T v) {
  return (w == T(0)) ? T(0) : w * v;
}

// A class test: unlike `x - x == 0` it cannot be rewritten by fp contraction
// (an fma of a product with its own negation returns the rounding error).
template <typename T>
__device__ __forceinline__ bool is_finite(T x) {
  return __builtin_isfinite(x);
}

}  // namespace lcm
