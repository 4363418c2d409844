// kg_common.cpp -- host-side helpers: error reporting, field geometry,
// binning heuristics and the exact squared-distance threshold.
#include "kg_common.hpp"


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

}  // namespace kg

extern "C" const char* kg_last_error(void) { return kg::t_last_error.c_str(); }
