// Error state of the C ABI (one message per calling thread).
#include <string>

#include "capi_common.h"

namespace {
thread_local std::string g_last_error;
}

int obte_fail(int code, const char* msg) {
  g_last_error = msg ? msg : "";
  return code;
}

extern "C" const char* obte_last_error(void) { return g_last_error.c_str(); }
