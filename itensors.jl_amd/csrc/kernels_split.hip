// Split-precision grouped GEMM for gfx950: FP64 operands emulated with
// integer slices on the INT8 matrix cores (Ozaki-style error-free splitting).
//
// 1. slice_kernel, once per operand: for every row r of every block (rows =
//    the GEMM's M modes for A, N modes for B; the K modes are the columns),
//    E_r = the smallest integer with max_k |x_rk| < 2^E_r, and
//        x_rk = 2^(E_r-6) * sum_{i=1..S} v_i,rk * 2^(-7(i-1)) + residue,
//    v_i in [-64, 64] (int8), |residue| <= 2^(E_r - 7S). The slices come
//    from exact FP64 steps (y = x*2^(6-E); v = rint(y); y = (y-v)*128).
//    The operand is read in place through its mode strides (no permutation
//    pass) and each slice plane is written row-major with K padded to a
//    multiple of 32 (zeros), so the GEMM reads 16-byte K chunks unchecked.
// 2. split_gemm_kernel: C = sum_{pairs} alpha_p * A_p B_p, where for every
//    pair the products v^A_i v^B_j with i + j <= S + 1 run on
//    v_mfma_i32_32x32x32_i8, summed per level L = i + j in INT32 -- exact,
//    since |sum| <= S * 64*64 * K < 2^31 for K <= 65536 (checked by the
//    planner) -- and recombined in FP64 after the pair's K loop:
//        C_mn += alpha * 2^(EA_m+EB_n-12) * sum_L 2^(-7(L-2)) S_L,mn.
//    The error is that of representing A and B with 7S bits per row/column
//    scale (about 2^-7S relative to the row maximum); no rounding occurs in
//    the products or the INT32 sums.
// Tiles: 256 threads = 4 wave64s (2x2), 64x64 per workgroup, 32x32 per wave;
// K steps of 32 double-buffered in LDS (slice planes, 16-byte chunks with a
// bank swizzle). The MFMA is issued as D = B^T A^T so that each lane's
// accumulators run down a column of C (contiguous column-major stores).
#include "itg_internal.h"

namespace itg {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;
constexpr int kRowsPerWG = 64;  // slice kernel rows per workgroup
constexpr int kNonFinite = 0x7fff;

__device__ inline int64_t find_slot(const int64_t* prefix, int64_t n, int64_t wg) {
  int64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= wg)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ inline int64_t mode_offset(uint32_t idx, int n, const int32_t* dim,
                                      const int32_t* str) {
  int64_t o = 0;
#pragma unroll
  for (int i = 0; i < kSliceModes; ++i) {
    if (i >= n) break;
    if (i + 1 < n) {
      const uint32_t q = idx / (uint32_t)dim[i];
      o += (int64_t)(idx - q * (uint32_t)dim[i]) * str[i];
      idx = q;
    } else {
      o += (int64_t)idx * str[i];
    }
  }
  return o;
}

// Magnitude used for the row scale: |x| (real) or max(|re|, |im|) (complex,
// whose two parts share one scale per row so that their slice products land
// on the same INT32 grid).
template <int CPLX>
__device__ inline double elem_mag(const double* x, int64_t o) {
  if (CPLX) {
    const double2 v = reinterpret_cast<const double2*>(x)[o];
    const double m = fmax(fabs(v.x), fabs(v.y));
    return (isfinite(v.x) && isfinite(v.y)) ? m : INFINITY;
  }
  const double v = x[o];
  return isfinite(v) ? fabs(v) : INFINITY;
}

// nsets: slice sets written per element -- 1 (real), 2 (re, im: A side of a
// complex product), 3 (re, im, -im: B side); set s of block b starts at
// planes + nsets*plane_off + s*S*R*Kpad.
template <int S, int CPLX>
__global__ void __launch_bounds__(kThreads)
    slice_kernel(const SliceBlock* __restrict__ blocks, const int64_t* __restrict__ prefix,
                 int64_t nblk, const double* __restrict__ src, int8_t* __restrict__ planes,
                 int32_t* __restrict__ exps, int nsets) {
  __shared__ double red[4][kRowsPerWG];
  __shared__ int32_t sexp[kRowsPerWG];
  const int64_t wg = blockIdx.x;
  const int64_t b = find_slot(prefix, nblk, wg);
  const SliceBlock& d = blocks[b];
  const int r0 = (int)(wg - prefix[b]) * kRowsPerWG;
  const int tid = threadIdx.x;
  const double* x = src + (CPLX ? 2 : 1) * d.off;

  // The source is contiguous either along the rows (A stored M-fastest) or
  // along K (A stored K-fastest, B in N form); threads are laid along the
  // contiguous direction so that loads coalesce.
  const bool kfast = d.nkm == 1 && d.kstr[0] == 1;

  // phase 1: per-row max |x| over K
  if (!kfast) {
    const int row = tid % kRowsPerWG, part = tid / kRowsPerWG;
    double m = 0.0;
    if (r0 + row < d.R) {
      const int64_t ro = mode_offset((uint32_t)(r0 + row), d.nrm, d.rdim, d.rstr);
      for (int k = part; k < d.K; k += 4)
        m = fmax(m, elem_mag<CPLX>(x, ro + mode_offset((uint32_t)k, d.nkm, d.kdim, d.kstr)));
    }
    red[part][row] = m;
  } else {
    // one wave per row at a time, lanes along K
    const int lane = tid & 63, wave = tid >> 6;
    for (int row = wave; row < kRowsPerWG; row += 4) {
      double m = 0.0;
      if (r0 + row < d.R) {
        const int64_t ro = mode_offset((uint32_t)(r0 + row), d.nrm, d.rdim, d.rstr);
        for (int k = lane; k < d.K; k += 64) m = fmax(m, elem_mag<CPLX>(x, ro + k));
      }
#pragma unroll
      for (int s = 32; s >= 1; s >>= 1) m = fmax(m, __shfl_xor(m, s));
      if (lane == 0) {
        red[0][row] = m;
        red[1][row] = red[2][row] = red[3][row] = 0.0;
      }
    }
  }
  __syncthreads();
  if (tid < kRowsPerWG) {
    const double m = fmax(fmax(red[0][tid], red[1][tid]), fmax(red[2][tid], red[3][tid]));
    int e = 0;
    if (!isfinite(m))
      e = kNonFinite;
    else if (m > 0.0)
      (void)frexp(m, &e);  // m = f*2^e, f in [0.5, 1): m < 2^e
    sexp[tid] = e;
    if (r0 + tid < d.R) exps[d.exp_off + r0 + tid] = e;
  }
  __syncthreads();

  // phase 2: slice 16 consecutive k of one row per item
  const int nchunk = d.Kpad / 16;
  const int items = kRowsPerWG * nchunk;
  const int64_t pstride = (int64_t)d.R * d.Kpad;
  for (int it = tid; it < items; it += kThreads) {
    // consecutive threads: consecutive rows (row-contiguous source) or
    // consecutive K chunks of one row (K-contiguous source)
    const int row = kfast ? it / nchunk : it % kRowsPerWG;
    const int ch = kfast ? it % nchunk : it / kRowsPerWG;
    const int gr = r0 + row;
    if (gr >= d.R) continue;
    const int e = sexp[row];
    const int64_t ro = mode_offset((uint32_t)gr, d.nrm, d.rdim, d.rstr);
    // K offsets of this chunk: step through the K modes instead of dividing
    int64_t ko = 0;
    int32_t kc[kSliceModes] = {0, 0, 0, 0};
    {
      uint32_t idx = (uint32_t)(ch * 16);
#pragma unroll
      for (int i = 0; i < kSliceModes; ++i) {
        if (i >= d.nkm) break;
        if (i + 1 < d.nkm) {
          const uint32_t q = idx / (uint32_t)d.kdim[i];
          kc[i] = (int32_t)(idx - q * (uint32_t)d.kdim[i]);
          idx = q;
        } else {
          kc[i] = (int32_t)idx;
        }
        ko += (int64_t)kc[i] * d.kstr[i];
      }
    }
    constexpr int NC = CPLX ? 3 : 1;  // re, im, -im
    uint32_t packed[NC][S][4];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int i = 0; i < S; ++i)
#pragma unroll
        for (int w = 0; w < 4; ++w) packed[c][i][w] = 0u;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int k = ch * 16 + j;
      double y = 0.0, yi = 0.0;
      if (k < d.K && e != kNonFinite) {
        if (CPLX) {
          const double2 v = reinterpret_cast<const double2*>(x)[ro + ko];
          y = ldexp(v.x, 6 - e);
          yi = ldexp(v.y, 6 - e);
        } else {
          y = ldexp(x[ro + ko], 6 - e);
        }
      }
      // advance the K coordinate by one (mode 0 fastest)
      ko += d.kstr[0];
      if (++kc[0] == d.kdim[0]) {
        kc[0] = 0;
        ko -= (int64_t)d.kdim[0] * d.kstr[0];
#pragma unroll
        for (int i = 1; i < kSliceModes; ++i) {
          if (i >= d.nkm) break;
          ko += d.kstr[i];
          if (++kc[i] < d.kdim[i]) break;
          kc[i] = 0;
          ko -= (int64_t)d.kdim[i] * d.kstr[i];
        }
      }
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const double v = rint(y);
        y = (y - v) * 128.0;
        packed[0][i][j >> 2] |= ((uint32_t)((int32_t)v & 0xff)) << (8 * (j & 3));
        if (CPLX) {
          const double w = rint(yi);
          yi = (yi - w) * 128.0;
          packed[1][i][j >> 2] |= ((uint32_t)((int32_t)w & 0xff)) << (8 * (j & 3));
          packed[NC - 1][i][j >> 2] |= ((uint32_t)((int32_t)(-w) & 0xff)) << (8 * (j & 3));
        }
      }
    }
    int8_t* dst = planes + (int64_t)nsets * d.plane_off + (int64_t)gr * d.Kpad + ch * 16;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c >= nsets) break;
#pragma unroll
      for (int i = 0; i < S; ++i)
        *reinterpret_cast<v4i*>(dst + (c * S + i) * pstride) = v4i{
            (int)packed[c][i][0], (int)packed[c][i][1], (int)packed[c][i][2], (int)packed[c][i][3]};
    }
  }
}

constexpr int BM = 64, BN = 64, BKB = 32;  // BKB: K bytes per LDS stage

// Complex data (CPLX = 1): one launch per output component `comp`; the
// component is a K-stacked real product over two halves, e.g.
//   Re C = [Re A | Im A] . [Re B ; -Im B],  Im C = [Re A | Im A] . [Im B ; Re B],
// with the slice sets of each half packed in `sets` (bits 0-1 / 2-3: A / B
// set of half 0, bits 4-5 / 6-7: half 1), accumulated into the same INT32
// levels (both halves share the row and column scales), and written to the
// interleaved complex output.
template <int S, int CPLX>
__global__ void __launch_bounds__(kThreads, CPLX ? 1 : 2)
    split_gemm_kernel(const SplitPair* __restrict__ pairs, const GemmGroup* __restrict__ groups,
                      const int2* __restrict__ tiles, const int8_t* __restrict__ planesA,
                      const int8_t* __restrict__ planesB, const int32_t* __restrict__ expA,
                      const int32_t* __restrict__ expB, double* __restrict__ Cg, double ga,
                      double beta, int comp, int sets, int nsetsA, int nsetsB) {
  constexpr int NL = S;  // levels 2..S+1
  constexpr int CH_A = S * BM * 2;  // 16-byte chunks per stage
  constexpr int CH_B = S * BN * 2;
  __shared__ v4i As[2][S][BM][2];
  __shared__ v4i Bs[2][S][BN][2];

  const int2 tt = tiles[blockIdx.x];
  if (tt.x < 0) return;
  const GemmGroup grp = groups[tt.x];
  const int tiles_m = (grp.M + BM - 1) / BM;
  const int tm = tt.y % tiles_m, tn = tt.y / tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 32, wn0 = (wave >> 1) * 32;
  const int l32 = lane & 31, lh = lane >> 5;

  // this lane's output rows/cols: m = m0+wm0+l32 ; n = n0+wn0+(r%4)+8*(r/4)+4*lh
  const int gm = m0 + wm0 + l32;
  double acc[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0;

  for (int p = grp.p0; p < grp.p1; ++p) {
    const SplitPair pr = pairs[p];
    const int kpad = pr.kpad;
    if (kpad <= 0) continue;
    const int8_t* pa = planesA;
    const int8_t* pb = planesB;
    v16i lev[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) lev[L] = v16i{0};

    // one 16-byte chunk = (slice plane, row, K half) of the A or B tile;
    // the CH_A + CH_B chunks of a stage are spread over the threads
    constexpr int NQ = (CH_A + CH_B + kThreads - 1) / kThreads;
    v4i rq[NQ];
    auto load = [&](int k0) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        int c = tid + q * kThreads;
        if (c >= CH_A + CH_B) continue;
        const bool isa = c < CH_A;
        if (!isa) c -= CH_A;
        const int half = c & 1, row = (c >> 1) % 64, pl = c / 128;
        const int g = (isa ? m0 : n0) + row;
        const bool in = g < (isa ? grp.M : grp.N);
        const int8_t* src = (isa ? pa + pl * pr.a_pstride : pb + pl * pr.b_pstride) +
                            (int64_t)g * kpad + k0 + half * 16;
        rq[q] = in ? *reinterpret_cast<const v4i*>(src) : v4i{0, 0, 0, 0};
      }
    };
    auto store = [&](int buf) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        int c = tid + q * kThreads;
        if (c >= CH_A + CH_B) continue;
        const bool isa = c < CH_A;
        if (!isa) c -= CH_A;
        const int half = c & 1, row = (c >> 1) % 64, pl = c / 128;
        if (isa)
          As[buf][pl][row][half ^ ((row >> 3) & 1)] = rq[q];
        else
          Bs[buf][pl][row][half ^ ((row >> 3) & 1)] = rq[q];
      }
    };

    for (int h = 0; h < (CPLX ? 2 : 1); ++h) {
    // slice sets (0 = re, 1 = im, 2 = -im) of this half, chosen by the host
    // for the output component and the conj flags
    if (CPLX) {
      const int seta = h == 0 ? (sets & 3) : ((sets >> 4) & 3);
      const int setb = h == 0 ? ((sets >> 2) & 3) : ((sets >> 6) & 3);
      pa = planesA + (int64_t)nsetsA * pr.a_planes + (int64_t)seta * S * pr.a_pstride;
      pb = planesB + (int64_t)nsetsB * pr.b_planes + (int64_t)setb * S * pr.b_pstride;
    } else {
      pa = planesA + pr.a_planes;
      pb = planesB + pr.b_planes;
    }
    // every wave is past the previous half's last barrier: buffer 0 is free
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < kpad; k0 += BKB) {
      const bool more = k0 + BKB < kpad;
      const int ra_row = wm0 + l32, rb_row = wn0 + l32;
      // all B slices stay in registers; A slices stream through one register
      // quad (keeps the kernel at two waves per SIMD)
      v4i bv[S];
#pragma unroll
      for (int j = 0; j < S; ++j) bv[j] = Bs[buf][j][rb_row][lh ^ ((rb_row >> 3) & 1)];
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const v4i av = As[buf][i][ra_row][lh ^ ((ra_row >> 3) & 1)];
#pragma unroll
        for (int j = 0; j < S - i; ++j)  // slice i of A with slice j of B, level i+j
          lev[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bv[j], av, lev[i + j], 0, 0, 0);
      }
      // next stage: fetched after the MFMA block so the staging registers are
      // not live across it; the other wave on this SIMD covers the latency
      if (more) {
        load(k0 + BKB);
        store(buf ^ 1);
      }
      __syncthreads();
      buf ^= 1;
    }
    }  // halves

    // FP64 recombination of the levels, scaled per row/column
    const int ea = (gm < grp.M) ? expA[pr.a_exp + gm] : 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int gn = n0 + wn0 + (r & 3) + 8 * (r >> 2) + 4 * lh;
      const int eb = (gn < grp.N) ? expB[pr.b_exp + gn] : 0;
      double h = (double)lev[NL - 1][r];
#pragma unroll
      for (int L = NL - 2; L >= 0; --L) h = fma(h, 0.0078125, (double)lev[L][r]);
      double v = ldexp(h, ea + eb - 12);
      if (ea == kNonFinite || eb == kNonFinite) v = __builtin_nan("");
      acc[r] = fma(ga * pr.alpha, v, acc[r]);
    }
  }

  if (gm >= grp.M) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int gn = n0 + wn0 + (r & 3) + 8 * (r >> 2) + 4 * lh;
    if (gn >= grp.N) continue;
    const int64_t o0 = grp.off_c + gm + (int64_t)gn * grp.ldc;
    const int64_t o = CPLX ? 2 * o0 + comp : o0;
    double out = acc[r];
    if (beta != 0.0) out = fma(beta, Cg[o], out);
    Cg[o] = out;
  }
}

}  // namespace

int split_slices(int prec) { return prec == 2 ? kSliceCount : 0; }

hipError_t launch_slice(hipStream_t s, int prec, int cplx, int nsets, const SliceBlock* blocks,
                        const int64_t* prefix, int64_t nblk, int64_t grid, const void* src,
                        void* planes, void* exps) {
  if (grid == 0) return hipSuccess;
  if (prec != 2) return hipErrorNotSupported;
  if (cplx)
    hipLaunchKernelGGL((slice_kernel<kSliceCount, 1>), dim3((unsigned)grid), dim3(kThreads), 0, s,
                       blocks, prefix, nblk, (const double*)src, (int8_t*)planes, (int32_t*)exps,
                       nsets);
  else
    hipLaunchKernelGGL((slice_kernel<kSliceCount, 0>), dim3((unsigned)grid), dim3(kThreads), 0, s,
                       blocks, prefix, nblk, (const double*)src, (int8_t*)planes, (int32_t*)exps,
                       1);
  return hipGetLastError();
}

int64_t slice_block_wgs(int32_t rows) { return (rows + kRowsPerWG - 1) / kRowsPerWG; }

hipError_t launch_split_gemm(hipStream_t s, int prec, int cplx, int comp, int sets,
                             const SplitPair* pairs, const GemmGroup* groups, const int2* tiles,
                             int64_t grid, const void* planesA, const void* planesB,
                             const void* expA, const void* expB, void* C, double ga,
                             double beta) {
  if (grid == 0) return hipSuccess;
  if (prec != 2) return hipErrorNotSupported;
  if (cplx)
    hipLaunchKernelGGL((split_gemm_kernel<kSliceCount, 1>), dim3((unsigned)grid), dim3(kThreads),
                       0, s, pairs, groups, tiles, (const int8_t*)planesA, (const int8_t*)planesB,
                       (const int32_t*)expA, (const int32_t*)expB, (double*)C, ga, beta, comp,
                       sets, 3, 3);
  else
    hipLaunchKernelGGL((split_gemm_kernel<kSliceCount, 0>), dim3((unsigned)grid), dim3(kThreads),
                       0, s, pairs, groups, tiles, (const int8_t*)planesA, (const int8_t*)planesB,
                       (const int32_t*)expA, (const int32_t*)expB, (double*)C, ga, beta, 0, 0, 1,
                       1);
  return hipGetLastError();
}

}  // namespace itg
