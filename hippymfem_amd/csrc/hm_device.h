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

// The contraction on the matrix cores, FP64 or FP32 (hm_mfma.hip).
int hm_element_matrices_mfma(const hm_problem* P, int i, int j, int64_t e0, int64_t e1,
                             int precision, void* out, void* stream);
// The sum-factorized contraction for tensor-product hexahedra (hm_sumfact.hip); returns
// HM_ERR_UNSUPPORTED when the problem has no tensor structure.
int hm_element_matrices_sumfact(const hm_problem* P, int i, int j, int64_t e0, int64_t e1,
                                int precision, void* out, void* stream);
// Two blocks of one shape from one pass of the sum-factorized kernel (hm_sumfact.hip).
int hm_element_matrices_sumfact_pair(const hm_problem* P, int i1, int j1, int i2, int j2, int64_t e0,
                                     int64_t e1, int precision, void* out1, void* out2, void* stream);
// Element vectors on tensor problems, sum-factorized point state (hm_sumfact.hip).
int hm_element_vectors_sumfact(const hm_problem* P, int i, int64_t e0, int64_t e1, double* out, void* stream);
// Fills n words of 4 or 8 bytes with zero on the stream (a kernel: safe under capture).
int hm_fill_zero(void* p, size_t n, int word_bytes, hipStream_t st);
// Sets *flag to 1 on the stream where one of the n words (fp64 for 8 bytes, fp32 for 4) is
// not finite; leaves it otherwise.  Nothing where flag is NULL.
int hm_flag_nonfinite(const void* p, size_t n, int word_bytes, int32_t* flag, hipStream_t st);
