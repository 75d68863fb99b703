// Low-precision element contraction (FP32, or BF16/FP16 products with FP32 accumulation and
// per-element scaling).  Geometry stays FP64.
#include <hip/hip_runtime.h>

#include "hm_common.h"
#include "hm_device.h"

int hm_element_matrices_lowp(const hm_problem* P, int i, int j, int64_t e0, int64_t e1,
                             int precision, void* out, void* stream) {
  (void)P, (void)i, (void)j, (void)e0, (void)e1, (void)out, (void)stream;
  HM_FAIL(HM_ERR_UNSUPPORTED, "precision %d is not built yet", precision);
}
