// Fiber-stationary direct contraction for gfx950.
//
// The MPO steps of an effective-Hamiltonian apply contract a large operand A
// (an environment-sized intermediate) with a tiny one B (an MPO tensor) over
// a handful of indices. Every element of A then feeds several output blocks
// (one per allowed B-free sector), so an output-stationary kernel that gives
// each output block its own workgroups fetches A once per output block. Here
// a thread owns one "fiber" -- a value of A's free multi-index -- and
// computes every output element that fiber feeds:
//
//   C_r[fiber, g] = beta*C_r + ga * sum_{pairs p of r} alpha_p
//                                  * sum_k A_p[fiber, k] B_p[k, g]
//
// for all output blocks r of the fiber's group and all B-free indices g, in
// plan order (deterministic, no atomics). Consecutive lanes take consecutive
// fibers along A's fastest free mode, so A reads and C writes coalesce; B is
// the same for every lane (broadcast). Repeated reads of a fiber's A values
// (one per output block) hit the caches: A crosses HBM once.
// Replaces, for this shape class, the per-pair permute + BLAS loop of
// NDTensors/src/blocksparse/contract_generic.jl:78-129.
#include "itg_internal.h"

namespace itg {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = kFibersPerWG / kThreads;

template <int CPLX, int NF>
__global__ void __launch_bounds__(kThreads)
    fiber_kernel(const FiberGroup* __restrict__ groups, const FiberOut* __restrict__ outs,
                 const FiberPair* __restrict__ pairs, const int32_t* __restrict__ koffs,
                 const int32_t* __restrict__ gofs_c, const int32_t* __restrict__ gofs_b,
                 const int32_t* __restrict__ wg_group, const double* __restrict__ A,
                 const double* __restrict__ B, double* __restrict__ C, unsigned flags,
                 double ga_re, double ga_im, double beta_re, double beta_im) {
  const int64_t wg = blockIdx.x;
  const FiberGroup& fg = groups[wg_group[wg]];
  const bool conja = (flags & 1u) != 0, conjb = (flags & 2u) != 0;
  const bool use_beta = (beta_re != 0.0 || beta_im != 0.0);
  const int64_t f0 = (wg - fg.wg0) * kFibersPerWG;
  // The thread's kPerThread fibers run in lockstep (they share every loop
  // bound), so each contracted index issues kPerThread independent loads.
  // Fibers past the group are clamped onto its last fiber and not stored.
  bool live[kPerThread];
  uint32_t c[kPerThread][NF];
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    const int64_t f = f0 + u * kThreads + threadIdx.x;
    live[u] = f < fg.nfib;
    uint32_t rem = (uint32_t)(live[u] ? f : fg.nfib - 1);
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      if (i + 1 < NF) {
        const uint32_t q = rem / (uint32_t)fg.fdim[i];
        c[u][i] = rem - q * (uint32_t)fg.fdim[i];
        rem = q;
      } else {
        c[u][i] = rem;
      }
    }
  }
  for (int r = fg.r0; r < fg.r1; ++r) {
    const FiberOut& o = outs[r];
    int64_t co[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      co[u] = o.off_c;
#pragma unroll
      for (int i = 0; i < NF; ++i) co[u] += (int64_t)c[u][i] * o.cstr[i];
    }
    for (int g = 0; g < o.ng; ++g) {
      double acc_re[kPerThread], acc_im[kPerThread];
#pragma unroll
      for (int u = 0; u < kPerThread; ++u) acc_re[u] = acc_im[u] = 0.0;
      for (int p = o.p0; p < o.p1; ++p) {
        const FiberPair& pr = pairs[p];
        int64_t ao[kPerThread];
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
          ao[u] = pr.off_a;
#pragma unroll
          for (int i = 0; i < NF; ++i) ao[u] += (int64_t)c[u][i] * pr.astr[i];
        }
        const int64_t bo = pr.off_b + gofs_b[pr.gb0 + g];
        const int32_t* ko = koffs + 2 * (int64_t)pr.kofs;
        double s_re[kPerThread], s_im[kPerThread];
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) s_re[u] = s_im[u] = 0.0;
        for (int k = 0; k < pr.nk; ++k) {
          const int32_t ka = ko[2 * k], kb = ko[2 * k + 1];
          if (CPLX) {
            const double2 y = reinterpret_cast<const double2*>(B)[bo + kb];
            const double yi = conjb ? -y.y : y.y;
            double2 x[kPerThread];
#pragma unroll
            for (int u = 0; u < kPerThread; ++u)
              x[u] = reinterpret_cast<const double2*>(A)[ao[u] + ka];
#pragma unroll
            for (int u = 0; u < kPerThread; ++u) {
              const double xi = conja ? -x[u].y : x[u].y;
              s_re[u] = fma(x[u].x, y.x, fma(-xi, yi, s_re[u]));
              s_im[u] = fma(x[u].x, yi, fma(xi, y.x, s_im[u]));
            }
          } else {
            const double y = B[bo + kb];
            double x[kPerThread];
#pragma unroll
            for (int u = 0; u < kPerThread; ++u) x[u] = A[ao[u] + ka];
#pragma unroll
            for (int u = 0; u < kPerThread; ++u) s_re[u] = fma(x[u], y, s_re[u]);
          }
        }
        const double al_re = ga_re * pr.alpha_re - ga_im * pr.alpha_im;
        const double al_im = ga_re * pr.alpha_im + ga_im * pr.alpha_re;
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
          if (CPLX) {
            acc_re[u] += al_re * s_re[u] - al_im * s_im[u];
            acc_im[u] += al_re * s_im[u] + al_im * s_re[u];
          } else {
            acc_re[u] = fma(al_re, s_re[u], acc_re[u]);
          }
        }
      }
      const int32_t go = gofs_c[o.g0 + g];
#pragma unroll
      for (int u = 0; u < kPerThread; ++u) {
        if (!live[u]) continue;
        const int64_t oo = co[u] + go;
        if (CPLX) {
          double2 out = make_double2(acc_re[u], acc_im[u]);
          if (use_beta) {
            const double2 cv = reinterpret_cast<double2*>(C)[oo];
            out.x += beta_re * cv.x - beta_im * cv.y;
            out.y += beta_re * cv.y + beta_im * cv.x;
          }
          reinterpret_cast<double2*>(C)[oo] = out;
        } else {
          double out = acc_re[u];
          if (use_beta) out = fma(beta_re, C[oo], out);
          C[oo] = out;
        }
      }
    }
  }
}

template <int CPLX>
hipError_t launch_c(hipStream_t s, int nf, dim3 grid, const FiberGroup* groups,
                    const FiberOut* outs, const FiberPair* pairs, const int32_t* koffs,
                    const int32_t* gofs_c, const int32_t* gofs_b, const int32_t* wg_group,
                    const void* A, const void* B, void* C, unsigned flags, double ar, double ai,
                    double br, double bi) {
#define ITG_NF_CASE(n)                                                                       \
  case n:                                                                                    \
    hipLaunchKernelGGL((fiber_kernel<CPLX, n>), grid, dim3(kThreads), 0, s, groups, outs,    \
                       pairs, koffs, gofs_c, gofs_b, wg_group, (const double*)A,             \
                       (const double*)B, (double*)C, flags, ar, ai, br, bi);                 \
    return hipGetLastError();
  switch (nf) {
    ITG_NF_CASE(1)
    ITG_NF_CASE(2)
    ITG_NF_CASE(3)
    ITG_NF_CASE(4)
    ITG_NF_CASE(5)
    ITG_NF_CASE(6)
    ITG_NF_CASE(7)
    ITG_NF_CASE(8)
    default:
      return hipErrorInvalidValue;
  }
#undef ITG_NF_CASE
}

}  // namespace

hipError_t launch_fiber(hipStream_t s, int cplx, int nf, const FiberGroup* groups,
                        const FiberOut* outs, const FiberPair* pairs, const int32_t* koffs,
                        const int32_t* gofs_c, const int32_t* gofs_b, const int32_t* wg_group,
                        int64_t total_wgs, const void* A, const void* B, void* C, unsigned flags,
                        double ga_re, double ga_im, double beta_re, double beta_im) {
  if (total_wgs == 0) return hipSuccess;
  dim3 grid((unsigned)total_wgs);
  if (cplx)
    return launch_c<1>(s, nf, grid, groups, outs, pairs, koffs, gofs_c, gofs_b, wg_group, A, B, C,
                       flags, ga_re, ga_im, beta_re, beta_im);
  return launch_c<0>(s, nf, grid, groups, outs, pairs, koffs, gofs_c, gofs_b, wg_group, A, B, C,
                     flags, ga_re, ga_im, beta_re, beta_im);
}

}  // namespace itg
