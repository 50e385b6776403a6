// Shared helpers of libptnhip.so: error reporting and complex element access.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/ptn_hip.h"

namespace ptn {

void set_error(const char *fmt, ...);

// Copy `bytes` of host data to device memory `dst`, ordered on stream `s`,
// through a ring of pinned staging buffers (the host data may be freed when
// the call returns).  Used for the per-instance descriptors of grouped launches.
int stage_to_device(hipStream_t s, const void *src, size_t bytes, void *dst);

#define PTN_HIP_CHECK(expr)                                                    \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::ptn::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,              \
                       hipGetErrorString(_e));                                 \
      return PTN_EHIP;                                                         \
    }                                                                          \
  } while (0)

#define PTN_LAUNCH_CHECK()                                                     \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess) {                                                    \
      ::ptn::set_error("%s:%d launch: %s", __FILE__, __LINE__,                 \
                       hipGetErrorString(_e));                                 \
      return PTN_EHIP;                                                         \
    }                                                                          \
  } while (0)

// complex element load/store with conversion to the arithmetic type
__device__ __forceinline__ void cload(const double2 *p, double &re, double &im) {
  double2 v = *p;
  re = v.x;
  im = v.y;
}
__device__ __forceinline__ void cload(const float2 *p, double &re, double &im) {
  float2 v = *p;
  re = v.x;
  im = v.y;
}
__device__ __forceinline__ void cload(const double2 *p, float &re, float &im) {
  double2 v = *p;
  re = (float)v.x;
  im = (float)v.y;
}
__device__ __forceinline__ void cload(const float2 *p, float &re, float &im) {
  float2 v = *p;
  re = v.x;
  im = v.y;
}
__device__ __forceinline__ void cstore(double2 *p, double re, double im) {
  *p = make_double2(re, im);
}
__device__ __forceinline__ void cstore(float2 *p, double re, double im) {
  *p = make_float2((float)re, (float)im);
}

}  // namespace ptn
