// Device-side helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "hm_common.h"

namespace hm {

// Reports a pending launch error as HM_ERR_HIP with the failing call's name.
inline int hip_check(const char* what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(err));
    return HM_ERR_HIP;
  }
  return HM_OK;
}

}  // namespace hm

// The low-precision contraction (hm_precision.hip).
int hm_element_matrices_lowp(const hm_problem* P, int i, int j, int64_t e0, int64_t e1,
                             int precision, void* out, void* stream);
