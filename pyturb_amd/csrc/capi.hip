// Error reporting and version of the C ABI.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>

#include "common.hpp"

namespace pt {

static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace pt

extern "C" const char* pt_last_error(void) { return pt::g_err; }

extern "C" int pt_abi_version(void) { return PT_ABI_VERSION; }
