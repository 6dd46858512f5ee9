// Shared helpers of the HIP kernels behind include/sumpyhip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "sumpyhip.h"

namespace sumpyhip {

int set_error(int code, const char *fmt, ...);

inline int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(SUMPYHIP_EHIP, "%s: launch failed: %s", what, hipGetErrorString(e));
  return SUMPYHIP_OK;
}

int num_cus();

struct cplx {
  double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re};
}
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.re*s, a.im*s}; }

// wave64 butterfly sum over groups of g consecutive lanes (g a power of two)
__device__ __forceinline__ double group_sum(double v, int g) {
  for (int o = g >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace sumpyhip

#define SUMPY_REQUIRE(cond, ...)                                   \
  do {                                                             \
    if (!(cond)) return ::sumpyhip::set_error(SUMPYHIP_EINVAL, __VA_ARGS__); \
  } while (0)
