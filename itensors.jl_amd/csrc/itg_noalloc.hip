// itg_bs_execute_noalloc: the block-sparse execute for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). Its workspace -- the permuted operands
// of the FP64 path, the int8 digit planes, row exponents and partial maxima
// of the split path -- lives in `scratch`, a fixed device array of
// ITG_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
// the process ends. Every call on one GPU shares that GPU's array: calls
// that overlap on one GPU (two streams, two host threads) race. The kernels
// write every workspace byte they read (slice planes including their K
// padding, exponents, partial maxima, permuted copies), so nothing is
// cleared between calls.
#include <mutex>

#include "../../include/itg.h"
#include "itg_plan.h"

namespace {

constexpr size_t kScratchBytes = ITG_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

// Host code must not pass `scratch` itself to a kernel (its host-side name
// is not a device address): noalloc_setup resolves each device's copy.
static __device__ __align__(256) char scratch[kScratchBytes];

namespace itg {

hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return hipSuccess;
  int prev = -1;
  hipError_t e = hipGetDevice(&prev);
  if (e != hipSuccess) return e;
  if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return e;
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess) return e;
  g_ptr[device] = p;
  return hipSuccess;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace itg

extern "C" itg_status itg_bs_execute_noalloc(itg_handle* h, itg_bs_plan* p, itg_dtype dtype,
                                             void* R, const void* T1, const void* T2,
                                             const double* alpha, const double* pair_alpha,
                                             itg_prec prec, unsigned flags) {
  if (!h || !p) return ITG_ERR_INVALID;
  void* ws = itg::noalloc_scratch(itg::handle_device(h));
  if (!ws)
    return itg::handle_fail(h, ITG_ERR_HIP,
                            "the scratch buffer of this device was not set up (itg_init failed "
                            "to resolve its address)");
  return itg::bs_execute_on(h, p, dtype, R, T1, T2, alpha, pair_alpha, prec, flags, ws,
                            kScratchBytes);
}
