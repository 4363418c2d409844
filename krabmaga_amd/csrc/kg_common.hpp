// kg_common.hpp -- host-side helpers shared by the C ABI implementations.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/krabgpu.h"
#include "kg_geom.hpp"
#include "kg_launch.hpp"

namespace kg {

// Records the message of the failing call for kg_last_error (per thread).
int set_error(int code, const char* fmt, ...);
void clear_error();

// Largest flock-kernel LDS capacity (staged agents) at or below `cap` whose
// dynamic plus static LDS fits the device's per-block limit.
int clamp_lds_cap(int cap);

// The gather rebinning after flock steps; KG_REBIN_GATHER=0 turns it off
// (scatter and bin sort every step), for comparison.
bool rebin_gather_enabled();

#define KG_HIP(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::kg::set_error(KG_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dmalloc(T** p, size_t n) {
  return hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
}

}  // namespace kg
