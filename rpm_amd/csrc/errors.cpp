// Thread-local error message storage for the C ABI.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "rpm_common.hpp"

namespace rpm {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace rpm

extern "C" int rpm_last_error(char* buffer, size_t len) {
  if (!buffer || len == 0) return RPM_E_INVALID;
  std::strncpy(buffer, rpm::g_last_error, len - 1);
  buffer[len - 1] = '\0';
  return RPM_OK;
}

extern "C" const char* rpm_version(void) { return "rpm_amd 0.1 (gfx950)"; }
