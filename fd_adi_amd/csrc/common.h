// Shared definitions for the fd_adi_amd HIP library (gfx950, FP64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/fdadi.h"

namespace fdadi {

// Thread-local last-error message; fdadi_last_error() returns it.
void set_error(const std::string &msg);

// Every entry point returns an int status; these helpers keep it uniform.
#define FDADI_TRY_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      fdadi::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));    \
      return FDADI_ERR_HIP;                                                   \
    }                                                                         \
  } while (0)

#define FDADI_REQUIRE(cond, code, msg)                                        \
  do {                                                                        \
    if (!(cond)) {                                                            \
      fdadi::set_error(msg);                                                  \
      return code;                                                            \
    }                                                                         \
  } while (0)

// Check the launch that was just issued.
#define FDADI_CHECK_LAUNCH() FDADI_TRY_HIP(hipGetLastError())

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

}  // namespace fdadi
