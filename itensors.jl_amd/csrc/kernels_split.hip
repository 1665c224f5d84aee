// Split-precision grouped GEMM (placeholder until the sliced kernels land).
#include "itg_internal.h"

namespace itg {

hipError_t launch_gemm_split(hipStream_t, int, int, int, int, const GemmPair*, const GemmGroup*,
                             const int32_t*, const int64_t*, int64_t, int64_t, const void*,
                             const void*, void*, double) {
  return hipErrorNotSupported;
}

}  // namespace itg
