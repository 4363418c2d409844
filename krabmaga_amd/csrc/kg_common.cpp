// kg_common.cpp -- host-side helpers shared by the C ABI implementations:
// per-thread error messages, the gather switch and the flock kernel's LDS
// capacity clamp.
#include "kg_common.hpp"

#include <cstdlib>


namespace kg {

static thread_local std::string t_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
  return code;
}

void clear_error() { t_last_error.clear(); }

bool rebin_gather_enabled() {
  static const bool v = [] {
    const char* e = std::getenv("KG_REBIN_GATHER");
    return !(e && e[0] == '0');
  }();
  return v;
}

// Keeps the flock kernel's LDS (dynamic staging + about 4 KB of static
// tables) within the device's per-block limit.
int clamp_lds_cap(int cap) {
  int dev = 0, max_lds = 65536, cu_lds = 160 * 1024;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 0)
      max_lds = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) ==
            hipSuccess &&
        v > 0)
      cu_lds = v;
  }
  // the flock kernel's work-groups per CU are set by their LDS (static
  // tables of about 4 KB plus the staged capacity): grow the capacity to the
  // largest that keeps the same count.  kStatic counts the static tables
  // with 2 KB to spare: at 1M boids a 28.7 KB capacity held four
  // work-groups per CU where 26.6 KB held five.
  constexpr size_t kStatic = 6144;
  const size_t per_wg = flock_lds_bytes(cap) + kStatic;
  if (per_wg <= (size_t)cu_lds) {
    const size_t share = (size_t)cu_lds / ((size_t)cu_lds / per_wg);
    while (flock_lds_bytes(cap + 64) + kStatic <= share) cap += 64;
  }
  const size_t budget = (size_t)max_lds - 8192;
  while (cap > 64 && flock_lds_bytes(cap) > budget) cap -= 64;
  return cap;
}

}  // namespace kg

extern "C" const char* kg_last_error(void) { return kg::t_last_error.c_str(); }
