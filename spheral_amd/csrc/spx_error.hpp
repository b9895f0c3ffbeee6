// spx_error.hpp -- error type and HIP status check shared by the host code.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/spx_abi.h"

namespace spx {

struct SpxError {
  int code;
  std::string msg;
};

[[noreturn]] inline void raise(int code, const std::string& msg) { throw SpxError{code, msg}; }

#define HIP_CHECK(expr)                                                                               \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) ::spx::raise(SPX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace spx
