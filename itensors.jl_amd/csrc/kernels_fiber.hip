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

#include <type_traits>

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

// Matrix form: out_o = sum_s a_s W[s][o] per fiber, W of the group in LDS.
// A workgroup builds W once and then streams kMVRounds rounds of one fiber
// per thread; the slots of the next round are loaded before the outputs of
// the current one are formed, so every wave keeps its A reads in flight
// behind its arithmetic and stores. Each A value is read once.
constexpr int kMVRounds = kFibersPerWGMV / kThreads;
static_assert(kMVRounds * kThreads == kFibersPerWGMV, "whole rounds per workgroup");

// SM: slot bound (4, 8 or 16) the slot loops are unrolled for; every group
// of the launch has nslot <= SM.
template <int CPLX, int NF, int SM>
__global__ void __launch_bounds__(kThreads)
    fiber_mv_kernel(const FiberMVGroup* __restrict__ groups, const FiberSlot* __restrict__ slots,
                    const FiberOutElem* __restrict__ outs, const FiberCoef* __restrict__ coefs,
                    const int32_t* __restrict__ wg_group, const double* __restrict__ A,
                    const double* __restrict__ B, double* __restrict__ C, unsigned flags,
                    double ga_re, double ga_im, double beta_re, double beta_im) {
  __shared__ double Wre[kFiberSlots * kFiberOuts];
  __shared__ double Wim[CPLX ? kFiberSlots * kFiberOuts : 1];
  // the group's slot and output descriptors, staged once per workgroup
  __shared__ int64_t s_off[kFiberSlots], o_off[kFiberOuts];
  __shared__ int32_t s_str[kFiberSlots][NF], o_str[kFiberOuts][NF];
  __shared__ uint32_t o_mask[kFiberOuts];
  const int64_t wg = blockIdx.x;
  const FiberMVGroup& fg = groups[wg_group[wg]];
  const int tid = threadIdx.x;
  const bool conja = (flags & 1u) != 0, conjb = (flags & 2u) != 0;
  const bool use_beta = (beta_re != 0.0 || beta_im != 0.0);
  const int nslot = fg.nslot, nout = fg.nout;

  // W = ga * alpha_p * B_p entries of the group, zero elsewhere
  for (int i = tid; i < kFiberSlots * kFiberOuts; i += kThreads) {
    Wre[i] = 0.0;
    if (CPLX) Wim[i] = 0.0;
  }
  if (tid < kFiberSlots) {
    const FiberSlot& sl = slots[fg.s0 + (tid < nslot ? tid : 0)];
    s_off[tid] = sl.off_a + sl.ka;
#pragma unroll
    for (int i = 0; i < NF; ++i) s_str[tid][i] = sl.astr[i];
  } else if (tid >= 64 && tid < 64 + kFiberOuts) {
    const int o = tid - 64;
    const FiberOutElem& oe = outs[fg.o0 + (o < nout ? o : 0)];
    o_off[o] = oe.off_c + oe.go;
    o_mask[o] = o < nout ? oe.mask : 0u;
#pragma unroll
    for (int i = 0; i < NF; ++i) o_str[o][i] = oe.cstr[i];
  }
  __syncthreads();
  for (int i = tid; i < fg.nw; i += kThreads) {
    const FiberCoef cf = coefs[fg.w0 + i];
    const double al_re = ga_re * cf.alpha_re - ga_im * cf.alpha_im;
    const double al_im = ga_re * cf.alpha_im + ga_im * cf.alpha_re;
    const int w = cf.out * kFiberSlots + cf.slot;  // output-major: one output's weights adjacent
    if (CPLX) {
      const double2 b = reinterpret_cast<const double2*>(B)[cf.off_b];
      const double bi = conjb ? -b.y : b.y;
      Wre[w] = al_re * b.x - al_im * bi;
      Wim[w] = al_re * bi + al_im * b.x;
    } else {
      Wre[w] = al_re * B[cf.off_b];
    }
  }
  __syncthreads();

  const int64_t fbase = (wg - fg.wg0) * kFibersPerWGMV;
  const int64_t flast = fg.nfib - 1;
  // rounds holding at least one fiber of the group (uniform)
  const int64_t left = fg.nfib - fbase;
  const int nrounds = (int)(left >= kFibersPerWGMV ? kMVRounds : (left + kThreads - 1) / kThreads);

  // coordinates of fiber f (clamped to the group's last fiber: a dead lane
  // reads valid memory and stores nothing)
  auto coords = [&](int64_t f, uint32_t (&c)[NF]) {
    uint32_t rem = (uint32_t)(f < flast ? f : flast);
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      if (i + 1 < NF) {
        const uint32_t q = rem / (uint32_t)fg.fdim[i];
        c[i] = rem - q * (uint32_t)fg.fdim[i];
        rem = q;
      } else {
        c[i] = rem;
      }
    }
  };
  // every slot of one fiber up to SM (slots past nslot repeat slot 0 and
  // are masked out below); fully unrolled so the values stay in registers and the
  // loads issue back to back
  auto load = [&](const uint32_t (&c)[NF], double (&xr)[SM], double (&xi)[CPLX ? SM : 1]) {
#pragma unroll
    for (int sl = 0; sl < SM; ++sl) {
      int64_t ao = s_off[sl];
#pragma unroll
      for (int i = 0; i < NF; ++i) ao += (int64_t)c[i] * s_str[sl][i];
      if (CPLX) {
        const double2 x = reinterpret_cast<const double2*>(A)[ao];
        xr[sl] = x.x;
        xi[CPLX ? sl : 0] = conja ? -x.y : x.y;
      } else {
        xr[sl] = A[ao];
      }
    }
  };

  uint32_t c[NF];
  double are[SM], aim[CPLX ? SM : 1];
  coords(fbase + tid, c);
  load(c, are, aim);
  for (int r = 0; r < nrounds; ++r) {
    // re-read the staged descriptors from LDS each round (broadcast reads)
    // instead of holding all of them in registers across the rounds
    asm volatile("" ::: "memory");
    const int64_t f = fbase + (int64_t)r * kThreads + tid;
    // next round's slots first (on the last round a clamped re-read of
    // the group's last fiber, which nothing uses)
    uint32_t cn[NF];
    double anr[SM], ani[CPLX ? SM : 1];
    coords(f + kThreads, cn);
    load(cn, anr, ani);
    const bool live = f < fg.nfib;
    // An unlinked (output, slot) pair holds 0 in LDS, so for finite slot
    // values the sum over all slots equals the sum over the linked ones bit
    // for bit (the accumulator starts at +0 and is never -0, and x * 0 =
    // +-0 leaves it unchanged); a wave holding any non-finite value masks,
    // so that an Inf or NaN reaches only the outputs it is linked to.
    bool fin = true;
#pragma unroll
    for (int sl = 0; sl < SM; ++sl) {
      fin = fin && isfinite(are[sl]);
      if (CPLX) fin = fin && isfinite(aim[CPLX ? sl : 0]);
    }
    auto outputs = [&](auto masked) {
      constexpr bool MASKED = decltype(masked)::value;
      for (int o = 0; o < nout; ++o) {
        const uint32_t m = MASKED ? o_mask[o] : ~0u;
        double acc_re = 0.0, acc_im = 0.0;
        // branch-free over all slots: an unlinked slot's product is formed
        // and discarded by a select (never added), so its value cannot leak
#pragma unroll
        for (int sl = 0; sl < SM; ++sl) {
          const bool lk = !MASKED || ((m >> sl) & 1u);
          const double wr = Wre[o * kFiberSlots + sl];
          if (CPLX) {
            const double wi = Wim[o * kFiberSlots + sl];
            const double tr = fma(are[sl], wr, fma(-aim[CPLX ? sl : 0], wi, acc_re));
            const double ti = fma(are[sl], wi, fma(aim[CPLX ? sl : 0], wr, acc_im));
            acc_re = lk ? tr : acc_re;
            acc_im = lk ? ti : acc_im;
          } else {
            const double t = fma(are[sl], wr, acc_re);
            acc_re = lk ? t : acc_re;
          }
        }
        if (live) {
          int64_t co = o_off[o];
#pragma unroll
          for (int i = 0; i < NF; ++i) co += (int64_t)c[i] * o_str[o][i];
          if (CPLX) {
            double2 out = make_double2(acc_re, acc_im);
            if (use_beta) {
              const double2 cv = reinterpret_cast<double2*>(C)[co];
              out.x += beta_re * cv.x - beta_im * cv.y;
              out.y += beta_re * cv.y + beta_im * cv.x;
            }
            reinterpret_cast<double2*>(C)[co] = out;
          } else {
            double out = acc_re;
            if (use_beta) out = fma(beta_re, C[co], out);
            C[co] = out;
          }
        }
      }
    };
    if (__all(fin))
      outputs(std::false_type{});
    else
      outputs(std::true_type{});
#pragma unroll
    for (int i = 0; i < NF; ++i) c[i] = cn[i];
#pragma unroll
    for (int sl = 0; sl < SM; ++sl) {
      are[sl] = anr[sl];
      if (CPLX) aim[CPLX ? sl : 0] = ani[CPLX ? sl : 0];
    }
  }
}

template <int CPLX, int SM>
hipError_t launch_mv(hipStream_t s, int nf, dim3 grid, const FiberMVGroup* groups,
                     const FiberSlot* slots, const FiberOutElem* outs, const FiberCoef* coefs,
                     const int32_t* wg_group, const void* A, const void* B, void* C,
                     unsigned flags, double ar, double ai, double br, double bi) {
#define ITG_NF_CASE(n)                                                                        \
  case n:                                                                                     \
    ITG_LAUNCH((fiber_mv_kernel<CPLX, n, SM>), grid, dim3(kThreads), 0, s, groups, slots,    \
               outs, coefs, wg_group, (const double*)A, (const double*)B, (double*)C, flags, \
               ar, ai, br, bi);                                                               \
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

template <int CPLX>
hipError_t launch_mv_sm(hipStream_t s, int nf, int smax, dim3 grid, const FiberMVGroup* groups,
                        const FiberSlot* slots, const FiberOutElem* outs,
                        const FiberCoef* coefs, const int32_t* wg_group, const void* A,
                        const void* B, void* C, unsigned flags, double ar, double ai, double br,
                        double bi) {
  switch (smax) {
    case 4:
      return launch_mv<CPLX, 4>(s, nf, grid, groups, slots, outs, coefs, wg_group, A, B, C,
                                flags, ar, ai, br, bi);
    case 8:
      return launch_mv<CPLX, 8>(s, nf, grid, groups, slots, outs, coefs, wg_group, A, B, C,
                                flags, ar, ai, br, bi);
    case kFiberSlots:
      return launch_mv<CPLX, kFiberSlots>(s, nf, grid, groups, slots, outs, coefs, wg_group, A,
                                          B, C, flags, ar, ai, br, bi);
    default:
      return hipErrorInvalidValue;
  }
}

// Fused pair of matrix-form steps (ChainGroup): per super-fiber, the x
// values are loaded once, y = W1 x is formed in registers and z = W2 y is
// stored; Y never reaches memory. W1 and W2 of the super-group sit in LDS.
// The arithmetic per y (per z) value is fiber_mv_kernel's, operation for
// operation, with unit scale, zero beta and no conjugation.
// The kernel is bound by memory latency. Real data with up to 8 x and y slots
// runs five waves per SIMD (96 VGPRs, a few spilled): 5 % faster than the
// four that its 106 VGPRs allow. Loading two rounds ahead instead (four
// waves, 126 VGPRs) is slower; the other forms spill too much at five.
template <int CPLX, int NF, int XM, int YM>
__global__ void __launch_bounds__(kThreads, (!CPLX && XM <= 8 && YM <= 8) ? 5 : 1)
    fiber_chain_kernel(const ChainGroup* __restrict__ groups, const FiberSlot* __restrict__ slots,
                       const FiberOutElem* __restrict__ zouts,
                       const uint32_t* __restrict__ ymask, const FiberCoef* __restrict__ coefs1,
                       const FiberCoef* __restrict__ coefs2,
                       const int32_t* __restrict__ wg_group, const double* __restrict__ X,
                       const double* __restrict__ B1, const double* __restrict__ B2,
                       double* __restrict__ Z) {
  __shared__ double W1re[kChainY * kChainX];
  __shared__ double W1im[CPLX ? kChainY * kChainX : 1];
  __shared__ double W2re[kChainZ * kChainY];
  __shared__ double W2im[CPLX ? kChainZ * kChainY : 1];
  __shared__ int64_t x_off[kChainX], z_off[kChainZ];
  __shared__ int32_t x_str[kChainX][NF], z_str[kChainZ][NF];
  __shared__ uint32_t y_mask[kChainY], z_mask[kChainZ];
  const int64_t wg = blockIdx.x;
  const ChainGroup& cg = groups[wg_group[wg]];
  const int tid = threadIdx.x;
  const int nx = cg.nx, ny = cg.ny, nz = cg.nz;

  for (int i = tid; i < kChainY * kChainX; i += kThreads) {
    W1re[i] = 0.0;
    if (CPLX) W1im[i] = 0.0;
  }
  for (int i = tid; i < kChainZ * kChainY; i += kThreads) {
    W2re[i] = 0.0;
    if (CPLX) W2im[i] = 0.0;
  }
  if (tid < kChainX) {
    const FiberSlot& sl = slots[cg.x0 + (tid < nx ? tid : 0)];
    x_off[tid] = sl.off_a + sl.ka;
#pragma unroll
    for (int i = 0; i < NF; ++i) x_str[tid][i] = sl.astr[i];
  } else if (tid >= 64 && tid < 64 + kChainZ) {
    const int o = tid - 64;
    const FiberOutElem& oe = zouts[cg.z0 + (o < nz ? o : 0)];
    z_off[o] = oe.off_c + oe.go;
    z_mask[o] = o < nz ? oe.mask : 0u;
#pragma unroll
    for (int i = 0; i < NF; ++i) z_str[o][i] = oe.cstr[i];
  } else if (tid >= 128 && tid < 128 + kChainY) {
    const int v = tid - 128;
    y_mask[v] = v < ny ? ymask[cg.y0 + v] : 0u;
  }
  __syncthreads();
  auto put = [&](const FiberCoef& cf, const double* Bp, double* Wr, double* Wi, int w) {
    // the scale factor is one: the products below are fiber_mv_kernel's
    const double ga_re = 1.0, ga_im = 0.0;
    const double al_re = ga_re * cf.alpha_re - ga_im * cf.alpha_im;
    const double al_im = ga_re * cf.alpha_im + ga_im * cf.alpha_re;
    if (CPLX) {
      const double2 b = reinterpret_cast<const double2*>(Bp)[cf.off_b];
      Wr[w] = al_re * b.x - al_im * b.y;
      Wi[w] = al_re * b.y + al_im * b.x;
    } else {
      Wr[w] = al_re * Bp[cf.off_b];
    }
  };
  for (int i = tid; i < cg.nw1; i += kThreads) {
    const FiberCoef cf = coefs1[cg.w10 + i];
    put(cf, B1, W1re, W1im, cf.out * kChainX + cf.slot);
  }
  for (int i = tid; i < cg.nw2; i += kThreads) {
    const FiberCoef cf = coefs2[cg.w20 + i];
    put(cf, B2, W2re, W2im, cf.out * kChainY + cf.slot);
  }
  __syncthreads();

  const int64_t fbase = (wg - cg.wg0) * kFibersPerWGChain;
  const int64_t flast = cg.nfib - 1;
  const int64_t left = cg.nfib - fbase;
  constexpr int kRounds = kFibersPerWGChain / kThreads;
  const int nrounds = (int)(left >= kFibersPerWGChain ? kRounds : (left + kThreads - 1) / kThreads);

  auto coords = [&](int64_t f, uint32_t (&c)[NF]) {
    uint32_t rem = (uint32_t)(f < flast ? f : flast);
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      if (i + 1 < NF) {
        const uint32_t q = rem / (uint32_t)cg.fdim[i];
        c[i] = rem - q * (uint32_t)cg.fdim[i];
        rem = q;
      } else {
        c[i] = rem;
      }
    }
  };
  auto load = [&](const uint32_t (&c)[NF], double (&xr)[XM], double (&xi)[CPLX ? XM : 1]) {
#pragma unroll
    for (int u = 0; u < XM; ++u) {
      int64_t ao = x_off[u];
#pragma unroll
      for (int i = 0; i < NF; ++i) ao += (int64_t)c[i] * x_str[u][i];
      if (CPLX) {
        const double2 x = reinterpret_cast<const double2*>(X)[ao];
        xr[u] = x.x;
        xi[CPLX ? u : 0] = x.y;
      } else {
        xr[u] = X[ao];
      }
    }
  };

  uint32_t c[NF];
  double xre[XM], xim[CPLX ? XM : 1];
  coords(fbase + tid, c);
  load(c, xre, xim);
  for (int r = 0; r < nrounds; ++r) {
    asm volatile("" ::: "memory");
    const int64_t f = fbase + (int64_t)r * kThreads + tid;
    uint32_t cn[NF];
    double xnr[XM], xni[CPLX ? XM : 1];
    coords(f + kThreads, cn);
    load(cn, xnr, xni);
    // y = W1 x, z = W2 y, each in slot order. A (slot, output) pair without
    // a coefficient holds 0 in LDS, so for finite values the unmasked sum
    // equals the masked one bit for bit: the accumulator starts at +0 and an
    // exact zero sum rounds to +0, so it is never -0, and adding x * 0 = +-0
    // leaves it unchanged. Inf or NaN times a missing coefficient would
    // not, so a wave holding any non-finite x (or y) takes the masked form.
    bool fin = true;
#pragma unroll
    for (int u = 0; u < XM; ++u) {
      fin = fin && isfinite(xre[u]);
      if (CPLX) fin = fin && isfinite(xim[CPLX ? u : 0]);
    }
    const bool fast_y = __all(fin);
    double yre[YM], yim[CPLX ? YM : 1];
    // MASKED = false: every term is added (the compile-time mask drops the
    // selects); true: only the pairs of the mask, as fiber_mv_kernel
    auto form_y = [&](auto masked) {
      constexpr bool MASKED = decltype(masked)::value;
#pragma unroll
      for (int v = 0; v < YM; ++v) {
        const uint32_t m = MASKED ? y_mask[v] : ~0u;
        double acc_re = 0.0, acc_im = 0.0;
#pragma unroll
        for (int u = 0; u < XM; ++u) {
          const bool lk = !MASKED || ((m >> u) & 1u);
          const double wr = W1re[v * kChainX + u];
          if (CPLX) {
            const double wi = W1im[v * kChainX + u];
            const double tr = fma(xre[u], wr, fma(-xim[CPLX ? u : 0], wi, acc_re));
            const double ti = fma(xre[u], wi, fma(xim[CPLX ? u : 0], wr, acc_im));
            acc_re = lk ? tr : acc_re;
            acc_im = lk ? ti : acc_im;
          } else {
            const double t = fma(xre[u], wr, acc_re);
            acc_re = lk ? t : acc_re;
          }
        }
        yre[v] = acc_re;
        if (CPLX) yim[CPLX ? v : 0] = acc_im;
      }
    };
    if (fast_y)
      form_y(std::false_type{});
    else
      form_y(std::true_type{});
    bool yfin = true;
#pragma unroll
    for (int v = 0; v < YM; ++v) {
      yfin = yfin && isfinite(yre[v]);
      if (CPLX) yfin = yfin && isfinite(yim[CPLX ? v : 0]);
    }
    const bool fast_z = fast_y && __all(yfin);
    const bool live = f < cg.nfib;
    auto form_z = [&](auto masked) {
      constexpr bool MASKED = decltype(masked)::value;
      for (int q = 0; q < nz; ++q) {
        const uint32_t m = MASKED ? z_mask[q] : ~0u;
        double acc_re = 0.0, acc_im = 0.0;
#pragma unroll
        for (int v = 0; v < YM; ++v) {
          const bool lk = !MASKED || ((m >> v) & 1u);
          const double wr = W2re[q * kChainY + v];
          if (CPLX) {
            const double wi = W2im[q * kChainY + v];
            const double tr = fma(yre[v], wr, fma(-yim[CPLX ? v : 0], wi, acc_re));
            const double ti = fma(yre[v], wi, fma(yim[CPLX ? v : 0], wr, acc_im));
            acc_re = lk ? tr : acc_re;
            acc_im = lk ? ti : acc_im;
          } else {
            const double t = fma(yre[v], wr, acc_re);
            acc_re = lk ? t : acc_re;
          }
        }
        if (live) {
          int64_t co = z_off[q];
#pragma unroll
          for (int i = 0; i < NF; ++i) co += (int64_t)c[i] * z_str[q][i];
          if (CPLX)
            reinterpret_cast<double2*>(Z)[co] = make_double2(acc_re, acc_im);
          else
            Z[co] = acc_re;
        }
      }
    };
    if (fast_z)
      form_z(std::false_type{});
    else
      form_z(std::true_type{});
#pragma unroll
    for (int i = 0; i < NF; ++i) c[i] = cn[i];
#pragma unroll
    for (int u = 0; u < XM; ++u) {
      xre[u] = xnr[u];
      if (CPLX) xim[CPLX ? u : 0] = xni[CPLX ? u : 0];
    }
  }
}

template <int CPLX, int XM, int YM>
hipError_t launch_chain(hipStream_t s, int nf, dim3 grid, const ChainGroup* groups,
                        const FiberSlot* slots, const FiberOutElem* zouts,
                        const uint32_t* ymask, const FiberCoef* coefs1, const FiberCoef* coefs2,
                        const int32_t* wg_group, const void* X, const void* B1, const void* B2,
                        void* Z) {
#define ITG_NF_CASE(n)                                                                        \
  case n:                                                                                     \
    ITG_LAUNCH((fiber_chain_kernel<CPLX, n, XM, YM>), grid, dim3(kThreads), 0, s, groups,    \
               slots, zouts, ymask, coefs1, coefs2, wg_group, (const double*)X,               \
               (const double*)B1, (const double*)B2, (double*)Z);                             \
    return hipGetLastError();
  switch (nf) {
    ITG_NF_CASE(1)
    ITG_NF_CASE(2)
    ITG_NF_CASE(3)
    ITG_NF_CASE(4)
    default:
      return hipErrorInvalidValue;
  }
#undef ITG_NF_CASE
}

template <int CPLX>
hipError_t launch_chain_xy(hipStream_t s, int nf, int xm, int ym, dim3 grid,
                           const ChainGroup* groups, const FiberSlot* slots,
                           const FiberOutElem* zouts, const uint32_t* ymask,
                           const FiberCoef* coefs1, const FiberCoef* coefs2,
                           const int32_t* wg_group, const void* X, const void* B1,
                           const void* B2, void* Z) {
  if (xm == 8 && ym == 8)
    return launch_chain<CPLX, 8, 8>(s, nf, grid, groups, slots, zouts, ymask, coefs1, coefs2,
                                    wg_group, X, B1, B2, Z);
  if (xm == kChainX && ym == kChainY)
    return launch_chain<CPLX, kChainX, kChainY>(s, nf, grid, groups, slots, zouts, ymask, coefs1,
                                                coefs2, wg_group, X, B1, B2, Z);
  if (xm == 8 && ym == kChainY)
    return launch_chain<CPLX, 8, kChainY>(s, nf, grid, groups, slots, zouts, ymask, coefs1,
                                          coefs2, wg_group, X, B1, B2, Z);
  if (xm == kChainX && ym == 8)
    return launch_chain<CPLX, kChainX, 8>(s, nf, grid, groups, slots, zouts, ymask, coefs1,
                                          coefs2, wg_group, X, B1, B2, Z);
  return hipErrorInvalidValue;
}

template <int CPLX>
hipError_t launch_c(hipStream_t s, int nf, dim3 grid, const FiberGroup* groups,
                    const FiberOut* outs, const FiberPair* pairs, const int32_t* koffs,
                    const int32_t* gofs_c, const int32_t* gofs_b, const int32_t* wg_group,
                    const void* A, const void* B, void* C, unsigned flags, double ar, double ai,
                    double br, double bi) {
#define ITG_NF_CASE(n)                                                                       \
  case n:                                                                                    \
    ITG_LAUNCH((fiber_kernel<CPLX, n>), grid, dim3(kThreads), 0, s, groups, outs,    \
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

hipError_t launch_fiber_mv(hipStream_t s, int cplx, int nf, int smax,
                           const FiberMVGroup* groups, const FiberSlot* slots,
                           const FiberOutElem* outs, const FiberCoef* coefs,
                           const int32_t* wg_group, int64_t total_wgs, const void* A,
                           const void* B, void* C, unsigned flags, double ga_re, double ga_im,
                           double beta_re, double beta_im) {
  if (total_wgs == 0) return hipSuccess;
  dim3 grid((unsigned)total_wgs);
  if (cplx)
    return launch_mv_sm<1>(s, nf, smax, grid, groups, slots, outs, coefs, wg_group, A, B, C,
                           flags, ga_re, ga_im, beta_re, beta_im);
  return launch_mv_sm<0>(s, nf, smax, grid, groups, slots, outs, coefs, wg_group, A, B, C, flags,
                         ga_re, ga_im, beta_re, beta_im);
}

hipError_t launch_fiber_chain(hipStream_t s, int cplx, int nf, int xm, int ym,
                              const ChainGroup* groups, const FiberSlot* slots,
                              const FiberOutElem* zouts, const uint32_t* ymask,
                              const FiberCoef* coefs1, const FiberCoef* coefs2,
                              const int32_t* wg_group, int64_t total_wgs, const void* X,
                              const void* B1, const void* B2, void* Z) {
  if (total_wgs == 0) return hipSuccess;
  dim3 grid((unsigned)total_wgs);
  if (cplx)
    return launch_chain_xy<1>(s, nf, xm, ym, grid, groups, slots, zouts, ymask, coefs1, coefs2,
                              wg_group, X, B1, B2, Z);
  return launch_chain_xy<0>(s, nf, xm, ym, grid, groups, slots, zouts, ymask, coefs1, coefs2,
                            wg_group, X, B1, B2, Z);
}

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
