// fc_runtime.hip — error channel and version query of libfluxcore_hip.so.
#include "fc_common.hpp"
#include <stdarg.h>
#include <stdio.h>

namespace {
thread_local char g_msg[512] = "";
}

extern "C" const char* fc_last_error(void) { return g_msg; }

extern "C" int fc_version(void) { return 1; }

namespace fc {

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(2, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
  return 0;
}

}  // namespace fc
