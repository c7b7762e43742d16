// fx_common.hpp -- error reporting shared by the host library.
//
// The reference prints and exits on any CUDA error (cuCheck,
// fluids5.0/helpers/common_defs.cpp:223-252).  Here every failing call
// records a message in a thread-local string and returns a negative status
// (FX_E* in fluidsx.h) to the caller.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "fluidsx.h"

namespace fx {

std::string& last_error();

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  last_error() = buf;
  return code;
}

inline int hip_fail(hipError_t e, const char* where, const char* what) {
  return fail(FX_EHIP, "%s: %s failed: %s", where, what, hipGetErrorString(e));
}

// A HIP event destroyed with its scope (error returns included).
struct Event {
  hipEvent_t e = nullptr;
  Event() = default;
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  hipError_t create() { return hipEventCreate(&e); }
};

}  // namespace fx

// Evaluates a HIP call; on failure records where/what and returns FX_EHIP.
#define FX_HIP(call, where)                                  \
  do {                                                       \
    hipError_t fx_e_ = (call);                               \
    if (fx_e_ != hipSuccess) return fx::hip_fail(fx_e_, where, #call); \
  } while (0)

// Propagates a non-zero status.
#define FX_TRY(expr)           \
  do {                         \
    int fx_s_ = (expr);        \
    if (fx_s_ != 0) return fx_s_; \
  } while (0)
