// Error convention shared by every C-ABI entry point of libopenbte_amd.
//
// Every `extern "C"` function returns an int: 0 on success, a negative code
// on failure (OBTE_E* in openbte_amd.h).  The message of the most recent
// failure on the calling thread is available from obte_last_error().  No C++
// exception crosses the ABI.
#pragma once

#include "openbte_amd.h"

// Record `msg` as the calling thread's last error and return `code`.
int obte_fail(int code, const char* msg);

// Record `who: <HIP error string>` and return OBTE_EHIP.
int obte_fail_hip(const char* who, int hip_error);
