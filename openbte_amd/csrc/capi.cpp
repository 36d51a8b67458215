// Error state of the C ABI (one message per calling thread).
#include <hip/hip_runtime_api.h>

#include <string>

#include "capi_common.h"

namespace {
thread_local std::string g_last_error;
}

int obte_fail(int code, const char* msg) {
  g_last_error = msg ? msg : "";
  return code;
}

int obte_fail_hip(const char* who, int hip_error) {
  g_last_error = std::string(who ? who : "") + ": " + hipGetErrorString(static_cast<hipError_t>(hip_error));
  return OBTE_EHIP;
}

extern "C" const char* obte_last_error(void) { return g_last_error.c_str(); }
