// errors.h -- error reporting shared by the C-ABI source files.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace nmx {
// record msg for nmx_last_error() (per thread) and return code
int set_error(int code, const std::string &msg);
}  // namespace nmx

#define NMX_HIPCHK(x)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess)                                                                    \
      return nmx::set_error(NMX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)
