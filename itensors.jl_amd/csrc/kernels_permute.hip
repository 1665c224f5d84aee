// Grouped tensor permutation for gfx950.
//
// dst(block b) = alpha * permutedims(src(block b), perm) + beta * dst(block b)
// for every block of a launch (one block for a dense tensor). The mode
// permutation is uniform across the launch; block dims differ.
//
// Two tilings, chosen on the host from the (merged) permutation:
//  kind 0: src mode 0 stays the fastest dst mode. Each workgroup streams
//          kElemsPerWG consecutive dst elements; reads and writes are both
//          contiguous along mode 0.
//  kind 1: src mode 0 and src mode perm[0] swap roles. 32x32 tiles staged
//          through LDS: read along src mode 0, write along dst mode 0, both
//          coalesced; the other modes are a batch index.
// Replaces NDTensors.permutedims! via Strided.jl (NDTensors/src/array/
// permutedims.jl:5-24); the alpha/beta form covers _contract_scalar_perm!
// (NDTensors/src/abstractarray/tensoralgebra/contract.jl:88-113).
#include "itg_internal.h"

namespace itg {

namespace {

constexpr int kThreads = 256;
constexpr int kElemsPerThread = 4;
constexpr int kElemsPerWG = kThreads * kElemsPerThread;
constexpr int kTile = 32;

struct PermArgs {
  int nm;
  int32_t perm[kMaxModes];
};

template <int CPLX>
struct Elem;
template <>
struct Elem<0> {
  using T = double;
};
template <>
struct Elem<1> {
  using T = double2;
};

__device__ inline double axpby(double a_re, double, double x, double b_re, double, double y,
                               bool use_beta) {
  double r = a_re * x;
  if (use_beta) r = fma(b_re, y, r);
  return r;
}
__device__ inline double2 axpby(double a_re, double a_im, double2 x, double b_re, double b_im,
                                double2 y, bool use_beta) {
  double2 r;
  r.x = a_re * x.x - a_im * x.y;
  r.y = a_re * x.y + a_im * x.x;
  if (use_beta) {
    r.x += b_re * y.x - b_im * y.y;
    r.y += b_re * y.y + b_im * y.x;
  }
  return r;
}

__device__ inline int64_t find_block(const int64_t* prefix, int64_t nblk, int64_t wg) {
  // largest b with prefix[b] <= wg
  int64_t lo = 0, hi = nblk - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= wg)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int CPLX>
__global__ void __launch_bounds__(kThreads)
    permute_inner_kernel(PermArgs pa, const PermBlock* __restrict__ blocks,
                         const int64_t* __restrict__ prefix, int64_t nblk,
                         const typename Elem<CPLX>::T* __restrict__ src,
                         typename Elem<CPLX>::T* __restrict__ dst, double a_re, double a_im,
                         double b_re, double b_im) {
  using T = typename Elem<CPLX>::T;
  const int64_t wg = blockIdx.x;
  const int64_t b = find_block(prefix, nblk, wg);
  const PermBlock blk = blocks[b];
  const int nm = pa.nm;
  // src strides and dst dims
  int64_t sstr[kMaxModes];
  int64_t ddim[kMaxModes];
  int64_t nelem = 1;
  {
    int64_t st = 1;
    for (int i = 0; i < nm; ++i) {
      sstr[i] = st;
      st *= blk.dims[i];
    }
    nelem = st;
    for (int i = 0; i < nm; ++i) ddim[i] = blk.dims[pa.perm[i]];
  }
  const bool use_beta = (b_re != 0.0 || b_im != 0.0);
  const int64_t e0 = (wg - prefix[b]) * kElemsPerWG;
  const int64_t d0 = ddim[0];
  for (int r = 0; r < kElemsPerThread; ++r) {
    const int64_t e = e0 + r * kThreads + threadIdx.x;
    if (e >= nelem) break;
    int64_t rem = e / d0;
    int64_t so = (e - rem * d0);  // mode 0 of dst == mode 0 of src (stride 1)
    for (int i = 1; i < nm; ++i) {
      const int64_t di = ddim[i];
      const int64_t q = rem / di;
      so += (rem - q * di) * sstr[pa.perm[i]];
      rem = q;
    }
    T x = src[blk.off_src + so];
    T y{};
    if (use_beta) y = dst[blk.off_dst + e];
    dst[blk.off_dst + e] = axpby(a_re, a_im, x, b_re, b_im, y, use_beta);
  }
}

template <int CPLX>
__global__ void __launch_bounds__(kThreads)
    permute_tile_kernel(PermArgs pa, const PermBlock* __restrict__ blocks,
                        const int64_t* __restrict__ prefix, int64_t nblk,
                        const typename Elem<CPLX>::T* __restrict__ src,
                        typename Elem<CPLX>::T* __restrict__ dst, double a_re, double a_im,
                        double b_re, double b_im) {
  using T = typename Elem<CPLX>::T;
  __shared__ T tile[kTile][kTile + 1];
  const int64_t wg = blockIdx.x;
  const int64_t b = find_block(prefix, nblk, wg);
  const PermBlock blk = blocks[b];
  const int nm = pa.nm;
  const int p = pa.perm[0];  // src mode that becomes dst's fastest mode
  int64_t sstr[kMaxModes];
  int64_t dstr_of_src[kMaxModes];  // dst stride of each src mode
  {
    int64_t st = 1;
    for (int i = 0; i < nm; ++i) {
      sstr[i] = st;
      st *= blk.dims[i];
    }
    int64_t dt = 1;
    for (int i = 0; i < nm; ++i) {
      dstr_of_src[pa.perm[i]] = dt;
      dt *= blk.dims[pa.perm[i]];
    }
  }
  const int64_t D0 = blk.dims[0];
  const int64_t Dp = blk.dims[p];
  const int64_t ti_n = (D0 + kTile - 1) / kTile;
  const int64_t tj_n = (Dp + kTile - 1) / kTile;
  int64_t t = wg - prefix[b];
  const int64_t ti = t % ti_n;
  t /= ti_n;
  const int64_t tj = t % tj_n;
  int64_t bidx = t / tj_n;
  // batch coordinates: every src mode except 0 and p, in src order
  int64_t sbase = 0, dbase = 0;
  for (int i = 1; i < nm; ++i) {
    if (i == p) continue;
    const int64_t di = blk.dims[i];
    const int64_t q = bidx / di;
    const int64_t c = bidx - q * di;
    sbase += c * sstr[i];
    dbase += c * dstr_of_src[i];
    bidx = q;
  }
  const int64_t i0 = ti * kTile, j0 = tj * kTile;
  const int lx = threadIdx.x % kTile;
  const int ly = threadIdx.x / kTile;  // 0..7
  // read: i along src mode 0 (contiguous)
  for (int r = 0; r < kTile; r += kThreads / kTile) {
    const int64_t i = i0 + lx, j = j0 + ly + r;
    if (i < D0 && j < Dp) tile[ly + r][lx] = src[blk.off_src + sbase + i + j * sstr[p]];
  }
  __syncthreads();
  const bool use_beta = (b_re != 0.0 || b_im != 0.0);
  const int64_t dstr0 = dstr_of_src[0];
  for (int r = 0; r < kTile; r += kThreads / kTile) {
    const int64_t j = j0 + lx, i = i0 + ly + r;
    if (i < D0 && j < Dp) {
      const int64_t o = blk.off_dst + dbase + j + i * dstr0;
      T y{};
      if (use_beta) y = dst[o];
      dst[o] = axpby(a_re, a_im, tile[lx][ly + r], b_re, b_im, y, use_beta);
    }
  }
}

template <int CPLX>
__global__ void __launch_bounds__(kThreads)
    scale_kernel(typename Elem<CPLX>::T* __restrict__ x, int64_t n, double b_re, double b_im) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  typename Elem<CPLX>::T y{};
  x[i] = axpby(b_re, b_im, x[i], 0.0, 0.0, y, false);
}

// Grouped strided copy: dst[off_dst + sum_m x_m dstr[m]] = src[off_src +
// sum_m x_m sstr[m]] over every multi-index x of each block. A workgroup
// takes kElemsPerWG consecutive elements of one block in the block's mode
// order (mode 0 fastest; the host puts the destination's unit-stride mode
// first so that stores coalesce).
template <int CPLX>
__global__ void __launch_bounds__(kThreads)
    copy_blocks_kernel(const CopyBlock* __restrict__ blocks, const int64_t* __restrict__ prefix,
                       int64_t nblk, int nm, const typename Elem<CPLX>::T* __restrict__ src,
                       typename Elem<CPLX>::T* __restrict__ dst) {
  const int64_t wg = blockIdx.x;
  const int64_t b = find_block(prefix, nblk, wg);
  const CopyBlock& d = blocks[b];
  const int64_t e0 = (wg - prefix[b]) * kElemsPerWG;
#pragma unroll
  for (int u = 0; u < kElemsPerThread; ++u) {
    const int64_t e = e0 + u * kThreads + threadIdx.x;
    if (e >= d.nelem) break;
    uint32_t rem = (uint32_t)e;
    int64_t so = d.off_src, dof = d.off_dst;
    for (int m = 0; m < nm; ++m) {
      uint32_t x = rem;
      if (m + 1 < nm) {
        const uint32_t q = rem / (uint32_t)d.dims[m];
        x = rem - q * (uint32_t)d.dims[m];
        rem = q;
      }
      so += (int64_t)x * d.sstr[m];
      dof += (int64_t)x * d.dstr[m];
    }
    dst[dof] = src[so];
  }
}

}  // namespace

hipError_t launch_scale(hipStream_t s, int cplx, void* x, int64_t n, double b_re, double b_im) {
  if (n <= 0) return hipSuccess;
  dim3 grid((unsigned)((n + kThreads - 1) / kThreads)), block(kThreads);
  if (cplx)
    ITG_LAUNCH(scale_kernel<1>, grid, block, 0, s, (double2*)x, n, b_re, b_im);
  else
    ITG_LAUNCH(scale_kernel<0>, grid, block, 0, s, (double*)x, n, b_re, b_im);
  return hipGetLastError();
}

int64_t permute_block_tiles(int nm, const int32_t* perm, const int32_t* dims, int mode_kind) {
  int64_t nelem = 1;
  for (int i = 0; i < nm; ++i) nelem *= dims[i];
  if (nelem == 0) return 0;
  if (mode_kind == 0) return (nelem + kElemsPerWG - 1) / kElemsPerWG;
  const int p = perm[0];
  const int64_t D0 = dims[0], Dp = dims[p];
  const int64_t nb = nelem / (D0 * Dp);
  return ((D0 + kTile - 1) / kTile) * ((Dp + kTile - 1) / kTile) * nb;
}

hipError_t launch_permute(hipStream_t s, int cplx, int nm, const int32_t* perm_host,
                          const PermBlock* blocks_dev, const int64_t* tile_prefix_dev,
                          int64_t nblk, int64_t total_tiles, int mode_kind, const void* src,
                          void* dst, double a_re, double a_im, double b_re, double b_im) {
  if (total_tiles == 0) return hipSuccess;
  PermArgs pa;
  pa.nm = nm;
  for (int i = 0; i < kMaxModes; ++i) pa.perm[i] = i < nm ? perm_host[i] : 0;
  dim3 grid((unsigned)total_tiles), block(kThreads);
  if (mode_kind == 0) {
    if (cplx)
      ITG_LAUNCH(permute_inner_kernel<1>, grid, block, 0, s, pa, blocks_dev,
                         tile_prefix_dev, nblk, (const double2*)src, (double2*)dst, a_re, a_im,
                         b_re, b_im);
    else
      ITG_LAUNCH(permute_inner_kernel<0>, grid, block, 0, s, pa, blocks_dev,
                         tile_prefix_dev, nblk, (const double*)src, (double*)dst, a_re, a_im,
                         b_re, b_im);
  } else {
    if (cplx)
      ITG_LAUNCH(permute_tile_kernel<1>, grid, block, 0, s, pa, blocks_dev,
                         tile_prefix_dev, nblk, (const double2*)src, (double2*)dst, a_re, a_im,
                         b_re, b_im);
    else
      ITG_LAUNCH(permute_tile_kernel<0>, grid, block, 0, s, pa, blocks_dev,
                         tile_prefix_dev, nblk, (const double*)src, (double*)dst, a_re, a_im,
                         b_re, b_im);
  }
  return hipGetLastError();
}

int64_t copy_block_wgs(int64_t nelem) { return (nelem + kElemsPerWG - 1) / kElemsPerWG; }

hipError_t launch_copy_blocks(hipStream_t s, int cplx, int nm, const CopyBlock* blocks_dev,
                              const int64_t* prefix_dev, int64_t nblk, int64_t total_wgs,
                              const void* src, void* dst) {
  if (total_wgs == 0) return hipSuccess;
  dim3 grid((unsigned)total_wgs), block(kThreads);
  if (cplx)
    ITG_LAUNCH(copy_blocks_kernel<1>, grid, block, 0, s, blocks_dev, prefix_dev, nblk, nm,
               (const double2*)src, (double2*)dst);
  else
    ITG_LAUNCH(copy_blocks_kernel<0>, grid, block, 0, s, blocks_dev, prefix_dev, nblk, nm,
               (const double*)src, (double*)dst);
  return hipGetLastError();
}

}  // namespace itg
