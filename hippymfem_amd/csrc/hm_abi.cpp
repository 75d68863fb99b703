// The error slot and version of the C ABI.
#include <cstdarg>
#include <cstdio>

#include "hm_common.h"

namespace hm {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

}  // namespace hm

extern "C" const char* hm_last_error(void) { return hm::g_error; }

extern "C" int hm_version(void) { return 1; }
