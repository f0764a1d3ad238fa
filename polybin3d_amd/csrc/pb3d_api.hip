// Library-level entry points: version, thread-local error message.
#include "pb3d_internal.h"

namespace pb3d {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace pb3d

extern "C" int pb3d_version(void) { return 10000; }
extern "C" const char* pb3d_last_error(void) { return pb3d::g_err.c_str(); }
