// Library-level C ABI: version and per-thread error messages.
#include "common.h"

namespace tqf {

static thread_local std::string g_error;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

void clear_error() { g_error.clear(); }

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    return set_error((int)e, "%s: %s", what, hipGetErrorString(e));
  }
  return TQF_OK;
}

}  // namespace tqf

extern "C" int tqf_version(void) { return 1; }

extern "C" const char* tqf_last_error(void) { return tqf::g_error.c_str(); }
