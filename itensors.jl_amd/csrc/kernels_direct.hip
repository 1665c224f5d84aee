// Direct (small-K) grouped tensor contraction for gfx950.
//
// For contractions whose contracted extent is small (the MPO applications of
// an effective Hamiltonian contract over (w, s) with K <= 10, or <= 2 per QN
// block pair), TTGT's permutations would move every byte several times for a
// handful of flops. This kernel instead reads A and B in place through their
// strides and writes each output element once:
//
//   C[c] = beta*C[c] (dense entry only) + sum_{pairs p of c's block}
//          alpha_p * sum_k A_p[a(c) + ka_p(k)] * B_p[b(c) + kb_p(k)]
//
// One workgroup covers kDirectElemsPerWG consecutive elements of one output
// block (column-major), so writes are contiguous and reads are contiguous
// along any output mode that is also a fastest-varying mode of A or B. The
// pairs of a block are summed in plan order: deterministic, no atomics.
// The number of output modes is a template parameter so that the per-mode
// coordinates stay in registers; index arithmetic is 32-bit (blocks hold
// fewer than 2^31 elements, which the planner checks).
// Replaces the per-pair permute + BLAS loop of contract_generic.jl:78-129 for
// this shape class.
#include "itg_internal.h"

namespace itg {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = kDirectElemsPerWG / kThreads;

template <int CPLX, int NR>
__global__ void __launch_bounds__(kThreads)
    direct_kernel(const DirectPair* __restrict__ pairs, const DirectGroup* __restrict__ groups,
                  const int32_t* __restrict__ koffs, const int32_t* __restrict__ wg_group,
                  const double* __restrict__ A, const double* __restrict__ B,
                  double* __restrict__ C, unsigned flags, double ga_re, double ga_im,
                  double beta_re, double beta_im) {
  const int64_t wg = blockIdx.x;
  const DirectGroup& grp = groups[wg_group[wg]];
  const uint32_t nelem = (uint32_t)grp.nelem;
  const uint32_t e0 = (uint32_t)(wg - grp.wg0) * kDirectElemsPerWG;
  const bool conja = (flags & 1u) != 0, conjb = (flags & 2u) != 0;
  const bool use_beta = (beta_re != 0.0 || beta_im != 0.0);
  uint32_t cdim[NR > 0 ? NR : 1];
#pragma unroll
  for (int i = 0; i < NR; ++i) cdim[i] = (uint32_t)grp.cdim[i];

  uint32_t coord[kPerThread][NR > 0 ? NR : 1];
  bool live[kPerThread];
  double acc_re[kPerThread], acc_im[kPerThread];
#pragma unroll
  for (int r = 0; r < kPerThread; ++r) {
    const uint32_t e = e0 + r * kThreads + threadIdx.x;
    live[r] = e < nelem;
    uint32_t rem = live[r] ? e : 0u;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      if (i + 1 < NR) {
        const uint32_t q = rem / cdim[i];
        coord[r][i] = rem - q * cdim[i];
        rem = q;
      } else {
        coord[r][i] = rem;
      }
    }
    acc_re[r] = 0.0;
    acc_im[r] = 0.0;
  }

  const int p1 = grp.p1;
  for (int p = grp.p0; p < p1; ++p) {
    const DirectPair& pr = pairs[p];
    const double al_re = ga_re * pr.alpha_re - ga_im * pr.alpha_im;
    const double al_im = ga_re * pr.alpha_im + ga_im * pr.alpha_re;
    int32_t cs_a[NR > 0 ? NR : 1], cs_b[NR > 0 ? NR : 1];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      cs_a[i] = pr.cstr_a[i];
      cs_b[i] = pr.cstr_b[i];
    }
    uint32_t oa[kPerThread], ob[kPerThread];
#pragma unroll
    for (int r = 0; r < kPerThread; ++r) {
      uint32_t a = 0, b = 0;
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        a += coord[r][i] * (uint32_t)cs_a[i];
        b += coord[r][i] * (uint32_t)cs_b[i];
      }
      oa[r] = a;
      ob[r] = b;
    }
    const double* Ap = A + (CPLX ? 2 : 1) * pr.off_a;
    const double* Bp = B + (CPLX ? 2 : 1) * pr.off_b;
    double s_re[kPerThread], s_im[kPerThread];
#pragma unroll
    for (int r = 0; r < kPerThread; ++r) {
      s_re[r] = 0.0;
      s_im[r] = 0.0;
    }
    const int32_t* ko = koffs + 2 * (int64_t)pr.kofs;
    const int nk = pr.nk;
    // two contracted indices per trip: the loads of both are in flight
    // together
#pragma unroll 2
    for (int k = 0; k < nk; ++k) {
      const uint32_t ka = (uint32_t)ko[2 * k], kb = (uint32_t)ko[2 * k + 1];
#pragma unroll
      for (int r = 0; r < kPerThread; ++r) {
        if (!live[r]) continue;
        if (CPLX) {
          const double2 x = reinterpret_cast<const double2*>(Ap)[oa[r] + ka];
          const double2 y = reinterpret_cast<const double2*>(Bp)[ob[r] + kb];
          const double xi = conja ? -x.y : x.y;
          const double yi = conjb ? -y.y : y.y;
          s_re[r] = fma(x.x, y.x, fma(-xi, yi, s_re[r]));
          s_im[r] = fma(x.x, yi, fma(xi, y.x, s_im[r]));
        } else {
          s_re[r] = fma(Ap[oa[r] + ka], Bp[ob[r] + kb], s_re[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kPerThread; ++r) {
      if (CPLX) {
        acc_re[r] += al_re * s_re[r] - al_im * s_im[r];
        acc_im[r] += al_re * s_im[r] + al_im * s_re[r];
      } else {
        acc_re[r] = fma(al_re, s_re[r], acc_re[r]);
      }
    }
  }

#pragma unroll
  for (int r = 0; r < kPerThread; ++r) {
    if (!live[r]) continue;
    const int64_t o = grp.off_c + (int64_t)(e0 + r * kThreads + threadIdx.x);
    if (CPLX) {
      double2 out = make_double2(acc_re[r], acc_im[r]);
      if (use_beta) {
        const double2 c = reinterpret_cast<double2*>(C)[o];
        out.x += beta_re * c.x - beta_im * c.y;
        out.y += beta_re * c.y + beta_im * c.x;
      }
      reinterpret_cast<double2*>(C)[o] = out;
    } else {
      double out = acc_re[r];
      if (use_beta) out = fma(beta_re, C[o], out);
      C[o] = out;
    }
  }
}

template <int CPLX, int NR>
hipError_t launch_nr(hipStream_t s, dim3 grid, const DirectPair* pairs, const DirectGroup* groups,
                     const int32_t* koffs, const int32_t* wg_group, const void* A, const void* B,
                     void* C, unsigned flags, double ar, double ai, double br, double bi) {
  ITG_LAUNCH((direct_kernel<CPLX, NR>), grid, dim3(kThreads), 0, s, pairs, groups, koffs,
                     wg_group, (const double*)A, (const double*)B, (double*)C, flags, ar, ai, br,
                     bi);
  return hipGetLastError();
}

template <int CPLX>
hipError_t launch_c(hipStream_t s, int nr, dim3 grid, const DirectPair* pairs,
                    const DirectGroup* groups, const int32_t* koffs, const int32_t* wg_group,
                    const void* A, const void* B, void* C, unsigned flags, double ar, double ai,
                    double br, double bi) {
#define ITG_NR_CASE(n)                                                                      \
  case n:                                                                                   \
    return launch_nr<CPLX, n>(s, grid, pairs, groups, koffs, wg_group, A, B, C, flags, ar, \
                              ai, br, bi);
  switch (nr) {
    ITG_NR_CASE(0)
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

int64_t direct_group_tiles(const DirectGroup* g, int nr) {
  (void)nr;
  return ((int64_t)g->nelem + kDirectElemsPerWG - 1) / kDirectElemsPerWG;
}

hipError_t launch_direct(hipStream_t s, int cplx, int nr, const DirectPair* pairs,
                         const DirectGroup* groups, const int32_t* koffs,
                         const int32_t* wg_group, int64_t total_tiles, const void* A,
                         const void* B, void* C, unsigned flags, double ga_re, double ga_im,
                         double beta_re, double beta_im) {
  if (total_tiles == 0) return hipSuccess;
  dim3 grid((unsigned)total_tiles);
  if (cplx)
    return launch_c<1>(s, nr, grid, pairs, groups, koffs, wg_group, A, B, C, flags, ga_re,
                       ga_im, beta_re, beta_im);
  return launch_c<0>(s, nr, grid, pairs, groups, koffs, wg_group, A, B, C, flags, ga_re, ga_im,
                     beta_re, beta_im);
}

}  // namespace itg
