// Split-precision grouped GEMM (placeholder until the sliced kernels land).
#include "itg_internal.h"

namespace itg {

hipError_t launch_gemm_split(hipStream_t, int, int, int, int, const GemmPair*, const GemmGroup*,
                             const int2*, int64_t, const void*, const void*, void*, double,
                             double) {
  return hipErrorNotSupported;
}

}  // namespace itg
