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
// Float32 / ComplexF32: formed in FP64, rounded once on the store.
__device__ inline float axpby(double a_re, double a_im, float x, double b_re, double b_im,
                              float y, bool use_beta) {
  return (float)axpby(a_re, a_im, (double)x, b_re, b_im, (double)y, use_beta);
}
__device__ inline float2 axpby(double a_re, double a_im, float2 x, double b_re, double b_im,
                               float2 y, bool use_beta) {
  const double2 r = axpby(a_re, a_im, double2{x.x, x.y}, b_re, b_im, double2{y.x, y.y}, use_beta);
  return float2{(float)r.x, (float)r.y};
}
__device__ inline bool nonzero(double x) { return x != 0.0; }
__device__ inline bool nonzero(double2 x) { return x.x != 0.0 || x.y != 0.0; }
__device__ inline bool nonzero(float x) { return x != 0.0f; }
__device__ inline bool nonzero(float2 x) { return x.x != 0.0f || x.y != 0.0f; }

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

__device__ inline uint32_t fdiv(uint32_t x, uint32_t mul, uint32_t shr) {
  return (__umulhi(x, mul) + x) >> shr;
}

// Src offset of dst element e of a block (dst modes 0..NM-1, mode 0 fastest).
template <int NM>
__device__ inline int64_t src_offset(const PermBlock& blk, uint32_t e) {
  int64_t so = 0;
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    if (i + 1 < NM) {
      const uint32_t q = fdiv(e, blk.dmul[i], blk.dshr[i]);
      so += (int64_t)(e - q * blk.ddim[i]) * blk.sstr[i];
      e = q;
    } else {
      so += (int64_t)e * blk.sstr[i];
    }
  }
  return so;
}

// Block-level "some written value is not zero" flag.
__device__ inline void mark_nonzero(int32_t* nzflags, int64_t b, bool nz) {
  if (!nzflags) return;
  if (__syncthreads_or(nz ? 1 : 0) && threadIdx.x == 0) nzflags[b] = 1;
}

// Kind 0: src mode 0 stays the fastest dst mode. A workgroup streams
// kElemsPerWG consecutive dst elements of one block: the four loads of a
// thread are issued back to back (offsets clamped into the block), then the
// stores; reads and writes run along mode 0.
template <typename T, int NM>
__global__ void __launch_bounds__(kThreads)
    permute_inner_kernel(const PermBlock* __restrict__ blocks, const int64_t* __restrict__ prefix,
                         int64_t nblk, const T* __restrict__ src, T* __restrict__ dst,
                         double a_re, double a_im, double b_re, double b_im,
                         int32_t* __restrict__ nzflags) {
  const int64_t wg = blockIdx.x;
  const int64_t b = find_block(prefix, nblk, wg);
  const PermBlock& blk = blocks[b];
  const int64_t off_src = blk.off_src, off_dst = blk.off_dst;
  const uint32_t n = (uint32_t)blk.nelem;
  const bool has_src = off_src >= 0;
  const bool use_beta = (b_re != 0.0 || b_im != 0.0) && !blk.fresh;
  const uint32_t e0 = (uint32_t)(wg - prefix[b]) * kElemsPerWG + threadIdx.x;
  T x[kElemsPerThread], y[kElemsPerThread];
#pragma unroll
  for (int r = 0; r < kElemsPerThread; ++r) {
    const uint32_t e = min(e0 + r * kThreads, n - 1);
    x[r] = has_src ? src[off_src + src_offset<NM>(blk, e)] : T{};
    y[r] = use_beta ? dst[off_dst + e] : T{};
  }
  bool nz = false;
#pragma unroll
  for (int r = 0; r < kElemsPerThread; ++r) {
    const uint32_t e = e0 + r * kThreads;
    if (e < n) {
      const T v = axpby(a_re, a_im, x[r], b_re, b_im, y[r], use_beta);
      nz |= nonzero(v);
      dst[off_dst + e] = v;
    }
  }
  mark_nonzero(nzflags, b, nz);
}

// Kind 1: dst mode 0 (src mode p) and dst mode q (src mode 0) form 32x32
// tiles staged through LDS: read along src mode 0, written along dst mode 0,
// both coalesced; the other dst modes are a batch index.
template <typename T, int NM>
__global__ void __launch_bounds__(kThreads)
    permute_tile_kernel(const PermBlock* __restrict__ blocks, const int64_t* __restrict__ prefix,
                        int64_t nblk, int q, const T* __restrict__ src, T* __restrict__ dst,
                        double a_re, double a_im, double b_re, double b_im,
                        int32_t* __restrict__ nzflags) {
  __shared__ T tile[kTile][kTile + 1];
  const int64_t wg = blockIdx.x;
  const int64_t b = find_block(prefix, nblk, wg);
  const PermBlock& blk = blocks[b];
  // dst strides (column-major over ddim); dst mode q is src mode 0
  int64_t dstr[NM];
  {
    int64_t acc = 1;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
      dstr[i] = acc;
      acc *= blk.ddim[i];
    }
  }
  uint32_t Dq = 1, Mq = 1, Sq = 0;
  int64_t dstr_q = 1;
#pragma unroll
  for (int i = 1; i < NM; ++i)
    if (i == q) {
      Dq = blk.ddim[i];
      Mq = blk.dmul[i];
      Sq = blk.dshr[i];
      dstr_q = dstr[i];
    }
  (void)Mq;
  (void)Sq;
  const uint32_t D0 = blk.ddim[0];
  const int64_t sstr0 = blk.sstr[0];  // src stride of dst mode 0 (src mode p)
  const uint32_t tq_n = (Dq + kTile - 1) / kTile, t0_n = (D0 + kTile - 1) / kTile;
  uint32_t t = (uint32_t)(wg - prefix[b]);
  const uint32_t tq = t % tq_n;
  t /= tq_n;
  const uint32_t t0 = t % t0_n;
  uint32_t bidx = t / t0_n;
  // batch coordinates: every dst mode except 0 and q
  int64_t sbase = 0, dbase = 0;
#pragma unroll
  for (int i = 1; i < NM; ++i) {
    if (i == q) continue;
    const uint32_t qq = fdiv(bidx, blk.dmul[i], blk.dshr[i]);
    const uint32_t c = bidx - qq * blk.ddim[i];
    sbase += (int64_t)c * blk.sstr[i];
    dbase += (int64_t)c * dstr[i];
    bidx = qq;
  }
  const uint32_t iq0 = tq * kTile, j0 = t0 * kTile;
  const int lx = threadIdx.x % kTile;
  const int ly = threadIdx.x / kTile;  // 0..7
  const bool has_src = blk.off_src >= 0;
  // read: iq (dst mode q = src mode 0, contiguous in src) along lanes
  if (has_src) {
#pragma unroll
    for (int r = 0; r < kTile; r += kThreads / kTile) {
      const uint32_t iq = iq0 + lx, j = j0 + ly + r;
      if (iq < Dq && j < D0) tile[ly + r][lx] = src[blk.off_src + sbase + iq + j * sstr0];
    }
  }
  __syncthreads();
  const bool use_beta = (b_re != 0.0 || b_im != 0.0) && !blk.fresh;
  bool nz = false;
#pragma unroll
  for (int r = 0; r < kTile; r += kThreads / kTile) {
    const uint32_t j = j0 + lx, iq = iq0 + ly + r;
    if (iq < Dq && j < D0) {
      const int64_t o = blk.off_dst + dbase + j + iq * dstr_q;
      T y{};
      if (use_beta) y = dst[o];
      const T x = has_src ? tile[lx][ly + r] : T{};
      const T v = axpby(a_re, a_im, x, b_re, b_im, y, use_beta);
      nz |= nonzero(v);
      dst[o] = v;
    }
  }
  mark_nonzero(nzflags, b, nz);
}

__global__ void __launch_bounds__(kThreads) fill_i32_kernel(int32_t* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) x[i] = 0;
}

template <int CPLX>
__global__ void __launch_bounds__(kThreads)
    scale_kernel(typename Elem<CPLX>::T* __restrict__ x, int64_t n, double b_re, double b_im) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  typename Elem<CPLX>::T y{};
  x[i] = axpby(b_re, b_im, x[i], 0.0, 0.0, y, false);
}

// x64 = x32 elementwise (n scalars: complex data counts both parts).
__global__ void __launch_bounds__(kThreads)
    widen_kernel(const float* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}

// dst[i] = src[i] converted to a wider element type (real -> complex puts
// zero imaginary parts; Float32 -> Float64 is exact).
template <typename S, typename D>
__device__ inline D widen_elem(S x);
template <> __device__ inline double widen_elem<float, double>(float x) { return x; }
template <> __device__ inline float2 widen_elem<float, float2>(float x) { return float2{x, 0.0f}; }
template <> __device__ inline double2 widen_elem<float, double2>(float x) { return double2{x, 0.0}; }
template <> __device__ inline double2 widen_elem<double, double2>(double x) { return double2{x, 0.0}; }
template <> __device__ inline double2 widen_elem<float2, double2>(float2 x) {
  return double2{x.x, x.y};
}
template <typename S, typename D>
__global__ void __launch_bounds__(kThreads)
    convert_kernel(const S* __restrict__ x, D* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) y[i] = widen_elem<S, D>(x[i]);
}

// y32 = round(x64 + beta * y32), the sum in FP64.
template <int CPLX>
__global__ void __launch_bounds__(kThreads)
    narrow_kernel(const typename Elem<CPLX>::T* __restrict__ x, void* __restrict__ y, int64_t n,
                  double b_re, double b_im) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const bool use_beta = b_re != 0.0 || b_im != 0.0;
  if (CPLX) {
    float2* yp = (float2*)y;
    const double2 v = ((const double2*)x)[i];
    float2 old{};
    if (use_beta) old = yp[i];
    double2 r{v.x, v.y};
    if (use_beta) {
      r.x += b_re * (double)old.x - b_im * (double)old.y;
      r.y += b_re * (double)old.y + b_im * (double)old.x;
    }
    yp[i] = float2{(float)r.x, (float)r.y};
  } else {
    float* yp = (float*)y;
    double r = ((const double*)x)[i];
    if (use_beta) r = fma(b_re, (double)yp[i], r);
    yp[i] = (float)r;
  }
}

// y = alpha * (x + i*x2) + beta * y over complex elements; x2 may be null.
__global__ void __launch_bounds__(kThreads)
    axpby_join_kernel(const double2* __restrict__ x, const double2* __restrict__ x2,
                      double2* __restrict__ y, int64_t n, double a_re, double a_im, double b_re,
                      double b_im) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  double2 v = x[i];
  if (x2) {
    const double2 w = x2[i];
    v.x -= w.y;
    v.y += w.x;
  }
  const bool use_beta = b_re != 0.0 || b_im != 0.0;
  double2 old{};
  if (use_beta) old = y[i];
  y[i] = axpby(a_re, a_im, v, b_re, b_im, old, use_beta);
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

hipError_t launch_widen(hipStream_t s, int cplx, int64_t n, const void* x32, void* x64) {
  const int64_t m = cplx ? 2 * n : n;
  if (m <= 0) return hipSuccess;
  ITG_LAUNCH(widen_kernel, dim3((unsigned)((m + kThreads - 1) / kThreads)), kThreads, 0, s,
             (const float*)x32, (double*)x64, m);
  return hipGetLastError();
}

hipError_t launch_convert(hipStream_t s, int from, int to, int64_t n, const void* x, void* y) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((n + kThreads - 1) / kThreads));
#define ITG_CONV(F, T, S, D)                                                              \
  if (from == F && to == T) {                                                            \
    ITG_LAUNCH((convert_kernel<S, D>), grid, kThreads, 0, s, (const S*)x, (D*)y, n);     \
    return hipGetLastError();                                                            \
  }
  ITG_CONV(kEltF32, kEltF64, float, double)
  ITG_CONV(kEltF32, kEltC64, float, float2)
  ITG_CONV(kEltF32, kEltC128, float, double2)
  ITG_CONV(kEltF64, kEltC128, double, double2)
  ITG_CONV(kEltC64, kEltC128, float2, double2)
#undef ITG_CONV
  return hipErrorInvalidValue;
}

hipError_t launch_narrow(hipStream_t s, int cplx, int64_t n, const void* x64, void* y32,
                         double b_re, double b_im) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((n + kThreads - 1) / kThreads));
  if (cplx)
    ITG_LAUNCH(narrow_kernel<1>, grid, kThreads, 0, s, (const double2*)x64, y32, n, b_re, b_im);
  else
    ITG_LAUNCH(narrow_kernel<0>, grid, kThreads, 0, s, (const double*)x64, y32, n, b_re, b_im);
  return hipGetLastError();
}

hipError_t launch_axpby_join(hipStream_t s, int64_t n, double a_re, double a_im, const void* x,
                             const void* x2, double b_re, double b_im, void* y) {
  if (n <= 0) return hipSuccess;
  dim3 grid((unsigned)((n + kThreads - 1) / kThreads)), block(kThreads);
  ITG_LAUNCH(axpby_join_kernel, grid, block, 0, s, (const double2*)x, (const double2*)x2,
             (double2*)y, n, a_re, a_im, b_re, b_im);
  return hipGetLastError();
}

int64_t permute_block_tiles(int nm, int q, const uint32_t* ddim, int mode_kind) {
  int64_t nelem = 1;
  for (int i = 0; i < nm; ++i) nelem *= ddim[i];
  if (nelem == 0) return 0;
  if (mode_kind == 0) return (nelem + kElemsPerWG - 1) / kElemsPerWG;
  const int64_t D0 = ddim[0], Dq = ddim[q];
  const int64_t nb = nelem / (D0 * Dq);
  return ((D0 + kTile - 1) / kTile) * ((Dq + kTile - 1) / kTile) * nb;
}

namespace {
template <typename T, int NM>
void launch_perm_nm(hipStream_t s, int kind, int q, dim3 grid, const PermBlock* blocks,
                    const int64_t* prefix, int64_t nblk, const void* src, void* dst, double a_re,
                    double a_im, double b_re, double b_im, int32_t* nz) {
  if (kind == 0)
    ITG_LAUNCH((permute_inner_kernel<T, NM>), grid, kThreads, 0, s, blocks, prefix, nblk,
               (const T*)src, (T*)dst, a_re, a_im, b_re, b_im, nz);
  else
    ITG_LAUNCH((permute_tile_kernel<T, NM>), grid, kThreads, 0, s, blocks, prefix, nblk, q,
               (const T*)src, (T*)dst, a_re, a_im, b_re, b_im, nz);
}

template <typename T>
void launch_perm_t(hipStream_t s, int nm, int kind, int q, dim3 grid, const PermBlock* blocks,
                   const int64_t* prefix, int64_t nblk, const void* src, void* dst, double a_re,
                   double a_im, double b_re, double b_im, int32_t* nz) {
  switch (nm) {
#define ITG_PERM_NM(N)                                                                        \
  case N:                                                                                     \
    launch_perm_nm<T, N>(s, kind, q, grid, blocks, prefix, nblk, src, dst, a_re, a_im, b_re, \
                         b_im, nz);                                                           \
    break;
    ITG_PERM_NM(1)
    ITG_PERM_NM(2)
    ITG_PERM_NM(3)
    ITG_PERM_NM(4)
    ITG_PERM_NM(5)
    ITG_PERM_NM(6)
    ITG_PERM_NM(7)
    ITG_PERM_NM(8)
#undef ITG_PERM_NM
  }
}
static_assert(kMaxModes == 8, "permute kernels are instantiated for 1..8 modes");
}  // namespace

hipError_t launch_permute(hipStream_t s, int elt, int nm, const int32_t* perm_host,
                          const PermBlock* blocks_dev, const int64_t* tile_prefix_dev,
                          int64_t nblk, int64_t total_tiles, int mode_kind, const void* src,
                          void* dst, double a_re, double a_im, double b_re, double b_im,
                          int32_t* nzflags) {
  if (nm < 1 || nm > kMaxModes) return hipErrorInvalidValue;
  if (nzflags && nblk > 0) {
    ITG_LAUNCH(fill_i32_kernel, dim3((unsigned)((nblk + kThreads - 1) / kThreads)), kThreads, 0,
               s, nzflags, nblk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (total_tiles == 0) return hipSuccess;
  int q = 0;  // dst mode fed by src mode 0
  for (int i = 0; i < nm; ++i)
    if (perm_host[i] == 0) q = i;
  const dim3 grid((unsigned)total_tiles);
  switch (elt) {
    case kEltF64:
      launch_perm_t<double>(s, nm, mode_kind, q, grid, blocks_dev, tile_prefix_dev, nblk, src,
                            dst, a_re, a_im, b_re, b_im, nzflags);
      break;
    case kEltC128:
      launch_perm_t<double2>(s, nm, mode_kind, q, grid, blocks_dev, tile_prefix_dev, nblk, src,
                             dst, a_re, a_im, b_re, b_im, nzflags);
      break;
    case kEltF32:
      launch_perm_t<float>(s, nm, mode_kind, q, grid, blocks_dev, tile_prefix_dev, nblk, src,
                           dst, a_re, a_im, b_re, b_im, nzflags);
      break;
    case kEltC64:
      launch_perm_t<float2>(s, nm, mode_kind, q, grid, blocks_dev, tile_prefix_dev, nblk, src,
                            dst, a_re, a_im, b_re, b_im, nzflags);
      break;
    default:
      return hipErrorInvalidValue;
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
