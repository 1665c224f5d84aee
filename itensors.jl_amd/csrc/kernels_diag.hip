// Block-sparse x block-diagonal contraction (DiagOut / DiagPair).
//
// The reference contracts each (block, diagonal block) pair by converting the
// diagonal block to a dense one and calling the dense contract!
// (NDTensors/src/blocksparse/diagblocksparse.jl:635-675 with
// NDTensors/src/diag/tensoralgebra/contract.jl:103-160). Here no dense block
// is formed: a diagonal block is nonzero only where all of its coordinates
// agree, so each output element needs one diagonal value and one T1 value
// per pair (or, when the diagonal tensor is fully contracted, a sum along
// the diagonal). The kernel streams the output once; the bound is HBM.
#include "itg_internal.h"

namespace itg {
namespace {

constexpr int kThreads = 256;

// A thread forms kPerThread output elements (kThreads apart), with every
// load of all of them issued before the first is used: addresses come from
// clamped indices (elements past the block onto its last one, diagonal
// positions past the diagonal onto its last one) and the values are masked
// afterwards, so the loads are unconditional and in flight together.
template <int CPLX, int NR>
__global__ void __launch_bounds__(kThreads)
    diag_contract_kernel(const DiagOut* __restrict__ outs, const DiagPair* __restrict__ pairs,
                         const int32_t* __restrict__ wg_out, const double* __restrict__ T1,
                         const double* __restrict__ D, int uniform, double u_re, double u_im,
                         double a_re, double a_im, double* __restrict__ R) {
  constexpr int kPerThread = kDiagChunk / kThreads;
  const int64_t wg = blockIdx.x;
  const DiagOut& o = outs[wg_out[wg]];
  const int64_t base = (wg - o.wg0) * kDiagChunk;
  int32_t c[kPerThread][NR];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t e = min(base + (int64_t)k * kThreads + threadIdx.x, o.nelem - 1);
    uint32_t rem = (uint32_t)e;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      if (i + 1 < NR) {
        const uint32_t q = rem / (uint32_t)o.dims[i];
        c[k][i] = (int32_t)(rem - q * (uint32_t)o.dims[i]);
        rem = q;
      } else {
        c[k][i] = (int32_t)rem;
      }
    }
  }
  double acc_re[kPerThread], acc_im[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) acc_re[k] = acc_im[k] = 0.0;
  for (int p = o.p0; p < o.p1; ++p) {
    const DiagPair& q = pairs[p];
    if (q.dlen <= 0) continue;  // an empty diagonal block adds nothing
    int64_t b[kPerThread];
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      b[k] = q.off1;
#pragma unroll
      for (int i = 0; i < NR; ++i) b[k] += (int64_t)c[k][i] * q.str1[i];
    }
    // one diagonal value and one T1 value per element
    auto term = [&](int k, int j, bool on) {
      const int64_t t = b[k] + (int64_t)j * q.kstr;
      if (CPLX) {
        const double2 x = reinterpret_cast<const double2*>(T1)[t];
        double dr = u_re, di = u_im;
        if (!uniform) {
          const double2 d = reinterpret_cast<const double2*>(D)[q.offD + j];
          dr = d.x;
          di = d.y;
        }
        const double tr = fma(dr, x.x, fma(-di, x.y, acc_re[k]));
        const double ti = fma(dr, x.y, fma(di, x.x, acc_im[k]));
        acc_re[k] = on ? tr : acc_re[k];
        acc_im[k] = on ? ti : acc_im[k];
      } else {
        const double d = uniform ? u_re : D[q.offD + j];
        const double tr = fma(d, T1[t], acc_re[k]);
        acc_re[k] = on ? tr : acc_re[k];
      }
    };
    if (q.u0 >= 0) {
      // the element's coordinate on D's first open mode fixes the diagonal
      // position; D's other open modes must equal it
#pragma unroll
      for (int k = 0; k < kPerThread; ++k) {
        int j = 0;
#pragma unroll
        for (int i = 0; i < NR; ++i)
          if (i == q.u0) j = c[k][i];
        bool ok = j < q.dlen;
#pragma unroll
        for (int i = 0; i < NR; ++i)
          if ((q.umask >> i) & 1u) ok = ok && c[k][i] == j;
        term(k, min(j, q.dlen - 1), ok);
      }
    } else {
      for (int j = 0; j < q.dlen; ++j)
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) term(k, j, true);
    }
  }
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t e = base + (int64_t)k * kThreads + threadIdx.x;
    if (e >= o.nelem) break;
    if (CPLX) {
      reinterpret_cast<double2*>(R)[o.offR + e] =
          make_double2(a_re * acc_re[k] - a_im * acc_im[k], a_re * acc_im[k] + a_im * acc_re[k]);
    } else {
      R[o.offR + e] = a_re * acc_re[k];
    }
  }
}

template <int CPLX>
hipError_t launch_c(hipStream_t s, int nr, dim3 grid, const DiagOut* outs, const DiagPair* pairs,
                    const int32_t* wg_out, const void* T1, const void* D, int uniform,
                    double u_re, double u_im, double a_re, double a_im, void* R) {
#define ITG_NR_CASE(n)                                                                       \
  case n:                                                                                    \
    ITG_LAUNCH((diag_contract_kernel<CPLX, n>), grid, dim3(kThreads), 0, s, outs, pairs,    \
               wg_out, (const double*)T1, (const double*)D, uniform, u_re, u_im, a_re, a_im, \
               (double*)R);                                                                  \
    return hipGetLastError();
  switch (nr) {
    ITG_NR_CASE(1)
    ITG_NR_CASE(2)
    ITG_NR_CASE(3)
    ITG_NR_CASE(4)
    ITG_NR_CASE(5)
    ITG_NR_CASE(6)
    ITG_NR_CASE(7)
    ITG_NR_CASE(8)
    default:
      return hipErrorInvalidValue;
  }
#undef ITG_NR_CASE
}

}  // namespace

hipError_t launch_diag(hipStream_t s, int cplx, int nr, const DiagOut* outs,
                       const DiagPair* pairs, const int32_t* wg_out, int64_t total_wgs,
                       const void* T1, const void* D, int uniform, double u_re, double u_im,
                       double a_re, double a_im, void* R) {
  if (total_wgs == 0) return hipSuccess;
  dim3 grid((unsigned)total_wgs);
  if (cplx)
    return launch_c<1>(s, nr, grid, outs, pairs, wg_out, T1, D, uniform, u_re, u_im, a_re, a_im,
                       R);
  return launch_c<0>(s, nr, grid, outs, pairs, wg_out, T1, D, uniform, u_re, u_im, a_re, a_im, R);
}

}  // namespace itg
