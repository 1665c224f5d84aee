// Split-precision grouped GEMM for gfx950: FP64 operands emulated with
// integer slices on the INT8 matrix cores (Ozaki-style error-free splitting).
//
// 1. slice_kernel, once per operand: for every row r of every block (rows =
//    the GEMM's M modes for A, N modes for B; the K modes are the columns),
//    E_r = the smallest integer with max_k |x_rk| < 2^E_r, and
//        x_rk = 2^E_r / 127 * sum_{i=0..S-1} v_i,rk * 254^(-i) + residue,
//    v_i in [-127, 127] (int8), |residue| <= 2^E_r / 127 * 254^(1-S) / 2:
//    y = 127 * (x / 2^E) (|y| < 127; one rounding, 2^-53 relative), then
//    v = rint(y), y = (y - v) * 254 (|y| <= 127), S times. With S = 6 the
//    residue is 3.7e-15 of the row maximum (the 7-bit digits of a power-of-two
//    radix would need 7 digits, 28 products, for the same order).
//    The operand is read in place through its mode strides (no permutation
//    pass). Each slice plane holds K padded to a multiple of 32 (zeros) in
//    16-byte chunks, chunk-major: chunk c of row r at byte (c*R + r)*16, so
//    that 64 consecutive rows of one chunk are one contiguous 1 KB run for
//    both the writer and the GEMM's loads.
// 2. split_gemm_kernel: C = sum_{pairs} alpha_p * A_p B_p, where for every
//    pair the products v^A_i v^B_j with i + j <= S - 1 (21 of them) run on
//    v_mfma_i32_32x32x32_i8, summed per level L = i + j in INT32 -- exact,
//    since |sum| <= S * 127^2 * K < 2^31 for K <= 2 * kSplitKChunk (the
//    planner cuts longer pairs into K chunks) -- and recombined in FP64
//    after the pair's K loop:
//        C_mn += alpha * 2^(EA_m+EB_n) / 127^2 * sum_L 254^(-L) S_L,mn.
//    The error is that of the two representations plus the dropped levels
//    L >= S (same order), about 1e-13 relative Frobenius for N(0,1) data;
//    no rounding occurs in the products or the INT32 sums.
// Tiles: 256 threads = 4 wave64s (2x2), (64*WMB) x 64 per workgroup,
// (32*WMB) x 32 per wave; K steps of 32 (two chunks) double-buffered in LDS.
// Stages are fetched with global_load_lds_dwordx4 (HBM/L2 -> LDS without
// registers): one wave instruction moves the 64 rows of one (plane, chunk),
// issued a full K step ahead of the MFMAs that consume it. In LDS a stage is
// [plane][chunk][row] x 16 bytes, so a lane's MFMA operand (row l%32, chunk
// l/32) is read with consecutive lanes on consecutive 16-byte slots (no bank
// conflicts). The MFMA is issued as D = B^T A^T so that each lane's
// accumulators run down a column of C (contiguous column-major stores).
#include "itg_internal.h"

namespace itg {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;
constexpr double kInvRadix = 1.0 / kSliceRadix;               // rounded: 2^-53 relative
constexpr double kInvLead2 = 1.0 / (kSliceLead * kSliceLead);  // rounded: 2^-53 relative
constexpr int kNonFinite = 0x7fff;   // row exponent of a row holding Inf / NaN
constexpr int kZeroRow = -0x40000000;  // partial exponent of an all-zero K range
static_assert(kSliceRows == 64 && kSliceKPart % 64 == 0, "slice tiling");

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

// Exponent recorded for a magnitude m: kNonFinite for Inf (elem_mag maps
// NaN to Inf), kZeroRow for 0, else e with m = f 2^e, f in [0.5, 1), so
// m < 2^e. Monotonic in m, so the exponent of a maximum is the maximum of
// the exponents.
__device__ inline int mag_exponent(double m) {
  int e = kZeroRow;
  if (!isfinite(m))
    e = kNonFinite;
  else if (m > 0.0)
    (void)frexp(m, &e);
  return e;
}

// Offset of K multi-index k (K mode 0 fastest), advanced one element at a
// time without divisions.
struct KWalk {
  int64_t off;
  int32_t c[kSliceModes];
  __device__ void init(uint32_t k, const SliceBlock& d) {
    off = 0;
#pragma unroll
    for (int i = 0; i < kSliceModes; ++i) {
      c[i] = 0;
      if (i >= d.nkm) continue;
      if (i + 1 < d.nkm) {
        const uint32_t q = k / (uint32_t)d.kdim[i];
        c[i] = (int32_t)(k - q * (uint32_t)d.kdim[i]);
        k = q;
      } else {
        c[i] = (int32_t)k;
      }
      off += (int64_t)c[i] * d.kstr[i];
    }
  }
  __device__ void step(const SliceBlock& d) {
    off += d.kstr[0];
    if (++c[0] == d.kdim[0]) {
      c[0] = 0;
      off -= (int64_t)d.kdim[0] * d.kstr[0];
#pragma unroll
      for (int i = 1; i < kSliceModes; ++i) {
        if (i >= d.nkm) break;
        off += d.kstr[i];
        if (++c[i] < d.kdim[i]) break;
        c[i] = 0;
        off -= (int64_t)d.kdim[i] * d.kstr[i];
      }
    }
  }
};

// Workgroup -> (block, 64-row tile rt, K part kp) of the slicing grid.
struct SliceTile {
  int64_t b;
  int r0, kp;
};
__device__ inline SliceTile slice_tile(const SliceBlock* blocks, const int2* wgmap, int64_t wg) {
  SliceTile t;
  const int2 m = wgmap[wg];
  t.b = m.x;
  const int kparts = blocks[t.b].kparts;
  t.kp = m.y % kparts;
  t.r0 = (m.y / kparts) * kSliceRows;
  return t;
}

// srow[i] = element offset of row min(r0 + i, R - 1) of block d (threads
// 0..kSliceRows-1 write; the caller synchronises).
__device__ inline void load_row_offsets(const SliceBlock& d, int r0, int tid, int64_t* srow) {
  if (tid < kSliceRows)
    srow[tid] = mode_offset((uint32_t)min(r0 + tid, d.R - 1), d.nrm, d.rdim, d.rstr);
}

// Pass 1: for each row of a (64 rows) x (kSliceKPart) tile, the exponent of
// the largest magnitude in the tile's K range (kZeroRow if all are zero,
// kNonFinite if any is Inf/NaN), stored as partial kp of the row. The source
// is contiguous either along the rows (operand stored rows-fastest) or along
// K; threads run along the contiguous direction so that loads coalesce.
template <int CPLX>
__global__ void __launch_bounds__(kThreads)
    slice_max_kernel(const SliceBlock* __restrict__ blocks, const int2* __restrict__ wgmap,
                     SliceOperands ops) {
  __shared__ int ered[4][kSliceRows];  // partial exponents: [thread part][row]
  __shared__ int64_t srow[kSliceRows];
  const SliceTile t = slice_tile(blocks, wgmap, blockIdx.x);
  const SliceBlock& d = blocks[t.b];
  const int tid = threadIdx.x;
  const double* x = ops.src[d.side] + (CPLX ? 2 : 1) * d.off;
  int32_t* pmax = ops.pmax[d.side];
  const int k_lo = t.kp * kSliceKPart, k_hi = min(d.K, k_lo + kSliceKPart);
  const bool kfast = d.nkm == 1 && d.kstr[0] == 1;
  // Every load below is unconditional, from an offset clamped into the block
  // (rows past R onto row R-1, K past the range onto its first element), and
  // the value is dropped afterwards: conditional loads end up in branches of
  // their own and serialise.
  load_row_offsets(d, t.r0, tid, srow);
  __syncthreads();
  if (!kfast) {
    // lanes along the rows; thread part p walks its quarter of the K range,
    // 16 elements at a time: the 16 offsets first, then 16 loads in flight
    constexpr int kQ = kSliceKPart / 4, kU = 16;
    const int row = tid % kSliceRows, part = tid / kSliceRows;
    double m = 0.0;
    const int ka = k_lo + part * kQ, kb = min(k_hi, ka + kQ);
    if (t.r0 + row < d.R && ka < kb && d.nkm == 1) {
      // one K mode (the usual case once adjacent modes are merged): the K
      // offset of element k is k * kstr for the whole wave, so all kQ loads
      // go out before the first is used (k past the range re-reads the
      // range's last element, which leaves the maximum unchanged)
      const int64_t ro = srow[row], ks = d.kstr[0];
      double v[kQ];
#pragma unroll
      for (int u = 0; u < kQ; ++u)
        v[u] = elem_mag<CPLX>(x, ro + (int64_t)min(ka + u, kb - 1) * ks);
#pragma unroll
      for (int u = 0; u < kQ; ++u) m = fmax(m, v[u]);
    } else if (t.r0 + row < d.R && ka < kb) {
      const int64_t ro = srow[row];
      KWalk w;
      w.init((uint32_t)ka, d);
      const int64_t o0 = w.off;
      for (int k = ka; k < kb; k += kU) {
        int64_t o[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          o[u] = (k + u < kb) ? w.off : o0;
          w.step(d);
        }
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) v[u] = elem_mag<CPLX>(x, ro + o[u]);
#pragma unroll
        for (int u = 0; u < kU; ++u) m = fmax(m, (k + u < kb) ? v[u] : 0.0);
      }
    }
    ered[part][row] = mag_exponent(m);
  } else {
    // lanes along K (two elements each: the part is 128 wide); a wave takes
    // rows wave, wave+4, ... and issues the loads of all of them before
    // reducing
    static_assert(kSliceKPart == 128, "two K elements per lane");
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int kRowsW = kSliceRows / 4;
    const int k0 = k_lo + lane, k1 = k0 + 64;
    const int c0 = min(k0, k_hi - 1), c1 = min(k1, k_hi - 1);
    double m[kRowsW];
#pragma unroll
    for (int i = 0; i < kRowsW; ++i) {
      const int64_t ro = srow[wave + 4 * i];
      const double a = elem_mag<CPLX>(x, ro + c0);
      const double b = elem_mag<CPLX>(x, ro + c1);
      m[i] = fmax(k0 < k_hi ? a : 0.0, k1 < k_hi ? b : 0.0);
    }
    // The 16 rows' exponents are reduced over the 64 lanes by a butterfly
    // that halves the rows a lane holds at each of the strides 32, 16, 8, 4
    // (15 lane exchanges instead of 16 x 6) and then reduces the quads;
    // lane l ends with row (l >> 2) of the wave.
    int w[kRowsW];
#pragma unroll
    for (int i = 0; i < kRowsW; ++i) w[i] = mag_exponent(m[i]);
#pragma unroll
    for (int n = kRowsW, b = 32; n > 1; n >>= 1, b >>= 1) {
      const bool hi = (lane & b) != 0;
#pragma unroll
      for (int j = 0; j < n / 2; ++j) {
        const int send = hi ? w[j] : w[j + n / 2];
        const int keep = hi ? w[j + n / 2] : w[j];
        w[j] = max(keep, __shfl_xor(send, b));
      }
    }
    w[0] = max(w[0], __shfl_xor(w[0], 2));
    w[0] = max(w[0], __shfl_xor(w[0], 1));
    if ((lane & 3) == 0) {
      const int row = wave + 4 * (lane >> 2);
      ered[0][row] = w[0];
      ered[1][row] = ered[2][row] = ered[3][row] = kZeroRow;
    }
  }
  __syncthreads();
  if (tid < kSliceRows && t.r0 + tid < d.R)
    pmax[d.pmax_off + (int64_t)t.kp * d.R + t.r0 + tid] =
        max(max(ered[0][tid], ered[1][tid]), max(ered[2][tid], ered[3][tid]));
}

// Pass 2: row exponent E = max of the row's partials (0 for an all-zero row),
// then the S slices of every element of the tile.
// nsets: slice sets written per element -- 1 (real) or 3 (re, im, -im:
// complex); set s of block b starts at planes + nsets*plane_off + s*S*R*Kpad.
template <int S, int CPLX>
__global__ void __launch_bounds__(kThreads)
    slice_kernel(const SliceBlock* __restrict__ blocks, const int2* __restrict__ wgmap,
                 SliceOperands ops, int nsets) {
  __shared__ int32_t sexp[kSliceRows];
  __shared__ int64_t srow[kSliceRows];
  const SliceTile t = slice_tile(blocks, wgmap, blockIdx.x);
  const SliceBlock& d = blocks[t.b];
  const int r0 = t.r0;
  const int tid = threadIdx.x;
  const double* x = ops.src[d.side] + (CPLX ? 2 : 1) * d.off;
  int8_t* planes = ops.planes[d.side];
  int32_t* exps = ops.exps[d.side];
  const int32_t* pmax = ops.pmax[d.side];
  load_row_offsets(d, r0, tid, srow);
  if (tid < kSliceRows) {
    int e = kZeroRow;
    if (r0 + tid < d.R)
      for (int p = 0; p < d.kparts; ++p)
        e = max(e, pmax[d.pmax_off + (int64_t)p * d.R + r0 + tid]);
    if (e == kZeroRow) e = 0;
    sexp[tid] = e;
    if (t.kp == 0 && r0 + tid < d.R) exps[d.exp_off + r0 + tid] = e;
  }
  __syncthreads();

  // slice 16 consecutive k of one row per item; consecutive threads take
  // consecutive rows, so the 16-byte chunks they write are adjacent in the
  // chunk-major planes (a K-contiguous source is then read 128 bytes per
  // lane, which the L1 cache serves line by line)
  constexpr int kChunks = kSliceKPart / 16;
  const int ch_end = min(d.Kpad / 16, (t.kp + 1) * kChunks);
  const int64_t pstride = (int64_t)d.R * d.Kpad;
  // A thread has kItems items; for real data the loads of all of them go
  // out before the first item's digits are formed (complex data: one item
  // at a time, for registers). Items past the block (rows >= R, chunks past
  // the K range) load from clamped, valid addresses and store nothing.
  constexpr int kItems = kSliceRows * kChunks / kThreads;
  static_assert(kSliceRows * kChunks % kThreads == 0, "items per thread");
  constexpr int NB = CPLX ? 1 : kItems;
  constexpr int NC = CPLX ? 3 : 1;  // re, im, -im
  for (int i0 = 0; i0 < kItems; i0 += NB) {
    double xr[NB][16], xi[NB][CPLX ? 16 : 1];
    int eq[NB], grq[NB], chq[NB];
    bool lq[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int it = tid + (i0 + q) * kThreads;
      const int row = it % kSliceRows;
      const int ch0 = t.kp * kChunks + it / kSliceRows;
      const int gr = r0 + row;
      lq[q] = gr < d.R && ch0 < ch_end;
      const int ch = min(ch0, ch_end - 1);
      const int e = sexp[row];
      eq[q] = e;
      grq[q] = gr;
      chq[q] = ch;
      const int64_t ro = srow[row];
      KWalk kw;
      kw.init((uint32_t)(ch * 16), d);
      // the 16 offsets first (padding k >= K clamped onto the row's k = 0),
      // then the 16 loads back to back and unconditional
      int64_t off[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        off[j] = ro + (ch * 16 + j < d.K ? kw.off : 0);
        kw.step(d);
      }
      const bool ok = e != kNonFinite;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const bool in = ok && ch * 16 + j < d.K;
        if (CPLX) {
          const double2 v = reinterpret_cast<const double2*>(x)[off[j]];
          xr[q][j] = in ? v.x : 0.0;
          xi[q][CPLX ? j : 0] = in ? v.y : 0.0;
        } else {
          const double v = x[off[j]];
          xr[q][j] = in ? v : 0.0;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      if (!lq[q]) continue;
      const int e = eq[q];
      // one slice set at a time (re, then im and -im for complex data: the
      // digits of -im are those of im negated), so only one set's packed
      // digits are live: with all three at once the complex kernel needed a
      // full register file and ran one wave per SIMD
      int8_t* dst =
          planes + (int64_t)nsets * d.plane_off + ((int64_t)chq[q] * d.R + grq[q]) * 16;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (c >= nsets) break;
        uint32_t packed[S][4];
#pragma unroll
        for (int i = 0; i < S; ++i)
#pragma unroll
          for (int w = 0; w < 4; ++w) packed[i][w] = 0u;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          // x / 2^E is exact (|.| < 1), the scale by 127 rounds once
          double y = ldexp(c == 0 ? xr[q][j] : xi[q][CPLX ? j : 0], -e) * kSliceLead;
#pragma unroll
          for (int i = 0; i < S; ++i) {
            const double v = rint(y);
            y = (y - v) * kSliceRadix;
            const int32_t dg = c == 2 ? -(int32_t)v : (int32_t)v;
            packed[i][j >> 2] |= ((uint32_t)(dg & 0xff)) << (8 * (j & 3));
          }
        }
#pragma unroll
        for (int i = 0; i < S; ++i)
          *reinterpret_cast<v4i*>(dst + (c * S + i) * pstride) =
              v4i{(int)packed[i][0], (int)packed[i][1], (int)packed[i][2], (int)packed[i][3]};
      }
    }
  }
}

constexpr int BN = 64, BKB = 32;  // BKB: K bytes per LDS stage (two 16-byte chunks)

typedef __attribute__((address_space(3))) void* lds_ptr_t;
typedef __attribute__((address_space(1))) void* gbl_ptr_t;

// 16 bytes per lane from global memory straight into LDS; lane l's bytes land
// at lds_base + 16*l (lds_base is wave-uniform, passed through M0).
__device__ inline void load_lds16(const int8_t* g, v4i* lds_base) {
  __builtin_amdgcn_global_load_lds((gbl_ptr_t)g, (lds_ptr_t)lds_base, 16, 0, 0);
}

// Complex data (CPLX = 1): one launch per output component `comp`; the
// component is a K-stacked real product over two halves, e.g.
//   Re C = [Re A | Im A] . [Re B ; -Im B],  Im C = [Re A | Im A] . [Im B ; Re B],
// with the slice sets of each half packed in `sets` (bits 0-1 / 2-3: A / B
// set of half 0, bits 4-5 / 6-7: half 1), accumulated into the same INT32
// levels (both halves share the row and column scales), and written to the
// interleaved complex output.
// s_waitcnt vmcnt(N) (expcnt and lgkmcnt left at their maxima).
template <int N>
__device__ inline void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | (((N >> 4) & 3) << 14));
}

// Workgroup barrier for the stage ring. __syncthreads() carries a release
// fence, which waits for every outstanding load -- including the LDS fills
// issued for later stages that are meant to stay in flight. Here the caller
// has already waited (vmcnt) for the stage the next step reads; this waits
// for the wave's own LDS reads (so no stage is refilled under a reader),
// then synchronises. The empty asm statements keep the compiler from moving
// memory operations across the barrier.
__device__ inline void stage_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xF | (7 << 4) | (0 << 8) | (3 << 14));  // lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// s_waitcnt lgkmcnt(0) alone.
__device__ inline void wait_lgkmcnt0() {
  __builtin_amdgcn_s_waitcnt(0xF | (7 << 4) | (0 << 8) | (3 << 14));
}

// 16-byte LDS read that the compiler does not track. The compiler's wait
// insertion makes every LDS read it knows of wait for every outstanding
// global->LDS fill, including fills of stages nobody reads yet, which
// flattens a multi-stage ring. The caller waits itself: the stage was
// completed by an explicit vmcnt wait and barrier, and the destination
// registers are used only after an explicit wait_lgkmcnt0() that the caller
// pins with a sched_barrier.
__device__ inline v4i lds_read16_untracked(uint32_t addr) {
  v4i v;
  asm volatile("ds_read_b128 %0, %1" : "=v"(v) : "v"(addr) : "memory");
  return v;
}

// WM waves along M (the tile is (32*WM) x 64, 2*WM waves, one 32x32 block
// each). WM = 2: 64 x 64 tiles, three LDS stages of 24 KB (each fetched two
// K steps ahead), two workgroups per CU. WM = 4: 128 x 64 tiles, four LDS
// stages of 36 KB (144 KB, each fetched three K steps ahead), one workgroup
// (eight waves) per CU. Either way a SIMD runs two waves. The fills bound
// this kernel more than the matrix cores do: with its MFMAs taken out it
// still takes three quarters of its time, without its fills a little over
// half; a stage waits for its slowest load, which an L2 miss makes long, so
// the ring is as deep as the LDS allows.
template <int S, int CPLX, int WM>
__global__ void __launch_bounds__(128 * WM, WM == 2 ? 2 : 1)
    split_gemm_kernel(const SplitPair* __restrict__ pairs, const GemmGroup* __restrict__ groups,
                      const int2* __restrict__ tiles, const int8_t* __restrict__ planesA,
                      const int8_t* __restrict__ planesB, const int32_t* __restrict__ expA,
                      const int32_t* __restrict__ expB, double* __restrict__ Cg, double ga,
                      double beta, int comp, int sets, int nsetsA, int nsetsB, int aim) {
  constexpr int BM = 32 * WM;
  constexpr int NW = 2 * WM;                    // waves
  constexpr int NSTAGE = WM == 2 ? 3 : 4;
  constexpr int RB = BM / 64;                   // 64-row blocks of the A tile
  constexpr int NL = S;                         // levels 0..S-1
  constexpr int A_ST = S * 2 * BM;              // 16-byte LDS slots of one A stage
  constexpr int B_ST = S * 2 * BN;
  constexpr int NLD_A = S * 2 * RB;             // wave loads (64 rows x 16 bytes) per A stage
  constexpr int NLD = NLD_A + S * 2;
  constexpr int NLD_W = (NLD + NW - 1) / NW;    // most loads a wave issues per stage
  constexpr int PER = NLD / NW;                 // fewest loads a wave issues per stage
  // wait until the loads of every stage but the newest `n` have landed (a
  // wave issuing more than PER per stage waits a little longer than needed)
  auto wait_newer = [](int n) {
    if (n >= 2)
      wait_vmcnt<2 * PER>();
    else if (n == 1)
      wait_vmcnt<PER>();
    else
      wait_vmcnt<0>();
  };
  static_assert(NSTAGE - 2 <= 2, "wait_newer covers at most two newer stages");
  __shared__ v4i lds[NSTAGE * (A_ST + B_ST)];

  const int2 tt = tiles[blockIdx.x];
  if (tt.x < 0) return;
  const GemmGroup grp = groups[tt.x];
  const int tiles_m = (grp.M + BM - 1) / BM;
  const int tm = tt.y % tiles_m, tn = tt.y / tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm0 = (wave % WM) * 32, wn0 = (wave / WM) * 32;
  const int l32 = lane & 31, lh = lane >> 5;
  // B row fetched by this lane; rows past the block are clamped onto its
  // last row (they only feed outputs that are never stored)
  const int brow = min(n0 + lane, grp.N - 1);

  // this lane's outputs: m = m0+wm0+l32 ; n = n0+wn0+(r%4)+8*(r/4)+4*lh
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

    // the NLD wave loads of the stage at K byte k0, into LDS stage `st`;
    // wave w issues loads w, w+NW, ...
    auto issue = [&](int k0, int st) {
      v4i* sA = lds + st * (A_ST + B_ST);
      v4i* sB = sA + A_ST;
      const int c0 = k0 >> 4;
#pragma unroll
      for (int t = 0; t < NLD_W; ++t) {
        const int q = wave + NW * t;
        if (q >= NLD) break;
        if (q < NLD_A) {
          const int rb = q % RB, ch = (q / RB) & 1, pl = q / (2 * RB);
          const int row = min(m0 + rb * 64 + lane, grp.M - 1);
          load_lds16(pa + pl * pr.a_pstride + ((int64_t)(c0 + ch) * grp.M + row) * 16,
                     sA + (pl * 2 + ch) * BM + rb * 64);
        } else {
          const int r = q - NLD_A, ch = r & 1, pl = r >> 1;
          load_lds16(pb + pl * pr.b_pstride + ((int64_t)(c0 + ch) * grp.N + brow) * 16,
                     sB + (pl * 2 + ch) * BN);
        }
      }
    };

    // this lane's row and column exponents, fetched before the K loop so
    // that their latency hides under it (indices clamped into the block: a
    // row or column past the block only feeds outputs that are not stored)
    const int ea = expA[pr.a_exp + min(m0 + wm0 + l32, grp.M - 1)];
    int eb[16];
#pragma unroll
    for (int r = 0; r < 16; ++r)
      eb[r] = expB[pr.b_exp + min(n0 + wn0 + (r & 3) + 8 * (r >> 2) + 4 * lh, grp.N - 1)];

    for (int h = 0; h < (CPLX ? 2 : 1); ++h) {
    // slice sets (0 = re, 1 = im, 2 = -im) of this half, chosen by the host
    // for the output component and the conj flags
    if (CPLX) {
      const int seta = h == 0 ? (sets & 3) : ((sets >> 4) & 3);
      const int setb = h == 0 ? ((sets >> 2) & 3) : ((sets >> 6) & 3);
      pa = planesA + (int64_t)nsetsA * pr.a_planes + pr.a_koff + (int64_t)seta * S * pr.a_pstride;
      pb = planesB + (int64_t)nsetsB * pr.b_planes + pr.b_koff + (int64_t)setb * S * pr.b_pstride;
    } else {
      pa = planesA + pr.a_planes + pr.a_koff;
      pb = planesB + pr.b_planes + pr.b_koff;
    }
    // Stage ring: step s computes from stage s % NSTAGE and fetches step
    // s + NSTAGE - 1 into the stage step s - 1 used (every wave finished
    // reading it before the barrier that closed step s - 1). Every wave is
    // past the previous half's / pair's last barrier, so all stages are free.
    const int nsteps = kpad / BKB;
    issue(0, 0);
    int pre = 1;  // stages issued ahead of the first step
    for (; pre < NSTAGE - 1 && pre < nsteps; ++pre) issue(pre * BKB, pre);
    wait_newer(pre - 1);  // stage 0 landed
    stage_barrier();
    int st = 0;
    for (int s = 0; s < nsteps; ++s) {
      const v4i* sA = lds + st * (A_ST + B_ST);
      const v4i* sB = sA + A_ST;
      // the fill of a later stage first, then the fragments in the order the
      // MFMAs consume them (A slice 0, the B slices, A slices 1..): the first
      // MFMA waits for two LDS reads, not fourteen, so a wave's reads overlap
      // the other waves' MFMAs
      const int nxt = s + NSTAGE - 1;
      int nst = st + NSTAGE - 1;
      if (nst >= NSTAGE) nst -= NSTAGE;
      if (nxt < nsteps) issue(nxt * BKB, nst);
      v4i bv[S], av[S];
      av[0] = sA[lh * BM + wm0 + l32];
#pragma unroll
      for (int j = 0; j < S; ++j) bv[j] = sB[(j * 2 + lh) * BN + wn0 + l32];
#pragma unroll
      for (int i = 1; i < S; ++i) av[i] = sA[(i * 2 + lh) * BM + wm0 + l32];
#pragma unroll
      for (int i = 0; i < S; ++i)
#pragma unroll
        for (int j = 0; j < S - i; ++j)  // slice i of A with slice j of B, level i+j
          lev[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bv[j], av[i], lev[i + j], 0, 0, 0);
      // All 2S fragment reads are issued before the first MFMA, each into its
      // own registers, and waited for once. Left to itself the scheduler
      // reuses one register set for the A slices and waits for every read
      // with lgkmcnt(0) between MFMA groups, draining the matrix pipe.
      __builtin_amdgcn_sched_group_barrier(0x100, 2 * S, 0);       // DS reads
      __builtin_amdgcn_sched_group_barrier(0x008, S * (S + 1) / 2, 0);  // MFMAs
      // The MFMAs only touch registers, so without this fence the scheduler
      // sinks them below the wait for the next stage, which then serialises
      // the fetch with the math.
      __builtin_amdgcn_sched_barrier(0);
      // stage s + 1 must have landed; the stages issued after it may stay in
      // flight
      wait_newer(min(nsteps - 1, s + NSTAGE - 1) - (s + 1));
      stage_barrier();
      if (++st == NSTAGE) st = 0;
    }
    }  // halves

    // FP64 recombination of the levels, scaled per row/column
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      double h = (double)lev[NL - 1][r];
#pragma unroll
      for (int L = NL - 2; L >= 0; --L) h = fma(h, kInvRadix, (double)lev[L][r]);
      double v = ldexp(h * kInvLead2, ea + eb[r]);
      if (ea == kNonFinite || eb[r] == kNonFinite) v = __builtin_nan("");
      acc[r] = fma(ga * (aim ? pr.alpha_im : pr.alpha), v, acc[r]);
    }
  }

  const int gm = m0 + wm0 + l32;
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

// Real split GEMM with the A operand in registers: 64 x 64 tiles, two waves
// per workgroup, each wave 32 rows x 64 columns (both 32-column B blocks
// share its A fragments). A fragments come straight from the chunk-major
// planes into registers (a wave owns its rows, nothing to share), prefetched
// one K step ahead; B goes through LDS (three stages, filled two steps
// ahead). Per MFMA this moves half the LDS bytes of the 32x32-per-wave
// kernels, which the LDS-activity counters show are LDS-bound. 7 levels x 2
// blocks of INT32 accumulators need one wave per SIMD (two workgroups per
// CU).
template <int S>
__global__ void __launch_bounds__(128, 1)
    split_gemm_areg_kernel(const SplitPair* __restrict__ pairs,
                           const GemmGroup* __restrict__ groups, const int2* __restrict__ tiles,
                           const int8_t* planesA, const int8_t* __restrict__ planesB,
                           const int32_t* __restrict__ expA, const int32_t* __restrict__ expB,
                           double* __restrict__ Cg, double ga, double beta) {
  constexpr int BM = 64;
  constexpr int NL = S;
  constexpr int B_ST = S * 2 * BN;  // 16-byte LDS slots of one B stage
  constexpr int NLD = S * 2;        // wave loads (64 rows x 16 bytes) per B stage
  constexpr int NLD_W = NLD / 2;    // per wave (two waves)
  static_assert(NLD % 2 == 0, "B stage loads split evenly over the two waves");
  constexpr int NSTAGE = 3;
  __shared__ v4i lds[NSTAGE * B_ST];

  const int2 tt = tiles[blockIdx.x];
  if (tt.x < 0) return;
  const GemmGroup grp = groups[tt.x];
  const int tiles_m = (grp.M + BM - 1) / BM;
  const int tm = tt.y % tiles_m, tn = tt.y / tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l32 = lane & 31, lh = lane >> 5;
  const int arow = min(m0 + wave * 32 + l32, grp.M - 1);  // A row of this lane's fragment
  const int brow = min(n0 + lane, grp.N - 1);             // B row this lane fetches
  const uint32_t lds_base = (uint32_t)(uintptr_t)(lds_ptr_t)lds;

  // outputs: m = m0 + 32*wave + l32 ; n = n0 + 32*jb + (r%4) + 8*(r/4) + 4*lh
  double acc[2][16];
#pragma unroll
  for (int jb = 0; jb < 2; ++jb)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[jb][r] = 0.0;

  for (int p = grp.p0; p < grp.p1; ++p) {
    const SplitPair pr = pairs[p];
    const int kpad = pr.kpad;
    if (kpad <= 0) continue;
    const int8_t* pa = planesA + pr.a_planes + pr.a_koff;
    const int8_t* pb = planesB + pr.b_planes + pr.b_koff;
    v16i lev[2][NL];
#pragma unroll
    for (int jb = 0; jb < 2; ++jb)
#pragma unroll
      for (int L = 0; L < NL; ++L) lev[jb][L] = v16i{0};

    auto fill_b = [&](int k0, int st) {
      v4i* sB = lds + st * B_ST;
      const int c0 = k0 >> 4;
#pragma unroll
      for (int t = 0; t < NLD_W; ++t) {
        const int q = wave + 2 * t, ch = q & 1, pl = q >> 1;
        load_lds16(pb + pl * pr.b_pstride + ((int64_t)(c0 + ch) * grp.N + brow) * 16,
                   sB + (pl * 2 + ch) * BN);
      }
    };
    v4i av[S], an[S];
    auto load_a = [&](int k0) {  // A fragments of the step at K byte k0 -> an
      const int c = (k0 >> 4) + lh;
#pragma unroll
      for (int i = 0; i < S; ++i)
        an[i] = *reinterpret_cast<const v4i*>(pa + i * pr.a_pstride +
                                              ((int64_t)c * grp.M + arow) * 16);
    };

    // prologue: B stage 0, A step 0, B stage 1 (in that order, so that
    // waiting for all but the last NLD_W loads covers the first two). Past
    // the last step the prefetches re-fetch the last step (clamped) into a
    // stage / registers nobody reads, so every step issues the same loads
    // and waits for the same count.
    const int nsteps = kpad / BKB;
    const int klast = (nsteps - 1) * BKB;
    fill_b(0, 0);
    load_a(0);
    fill_b(min(BKB, klast), 1);
    wait_vmcnt<NLD_W>();
    stage_barrier();
#pragma unroll
    for (int i = 0; i < S; ++i) av[i] = an[i];
    int st = 0;
    for (int s = 0; s < nsteps; ++s) {
      // A of the next step first, then the B fill two steps ahead: the wait
      // at the end of the step then leaves only that fill in flight
      load_a(min((s + 1) * BKB, klast));
      int nst = st + 2;
      if (nst >= NSTAGE) nst -= NSTAGE;
      fill_b(min((s + 2) * BKB, klast), nst);
      // B fragments of this stage; the second block's reads complete under
      // the first block's MFMAs. Both prefetches stay ahead of the math.
      const uint32_t sb = lds_base + (uint32_t)(st * B_ST + lh * BN + l32) * 16u;
      v4i bv[2][S];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < S; ++j) bv[0][j] = lds_read16_untracked(sb + (j * 2 * BN) * 16);
      wait_lgkmcnt0();
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < S; ++j) bv[1][j] = lds_read16_untracked(sb + (j * 2 * BN + 32) * 16);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int jb = 0; jb < 2; ++jb) {
        if (jb == 1) {
          __builtin_amdgcn_sched_barrier(0);
          wait_lgkmcnt0();
          __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int i = 0; i < S; ++i)
#pragma unroll
          for (int j = 0; j < S - i; ++j)
            lev[jb][i + j] =
                __builtin_amdgcn_mfma_i32_32x32x32_i8(bv[jb][j], av[i], lev[jb][i + j], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
      wait_vmcnt<NLD_W>();
      stage_barrier();
#pragma unroll
      for (int i = 0; i < S; ++i) av[i] = an[i];
      if (++st == NSTAGE) st = 0;
    }
    // the last (unused) fill must land before the next pair's prologue
    // writes the stages again
    wait_vmcnt<0>();
    stage_barrier();

    // FP64 recombination of the levels, scaled per row/column
    const int gm = m0 + wave * 32 + l32;
    const int ea = (gm < grp.M) ? expA[pr.a_exp + gm] : 0;
#pragma unroll
    for (int jb = 0; jb < 2; ++jb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int gn = n0 + jb * 32 + (r & 3) + 8 * (r >> 2) + 4 * lh;
        const int eb = (gn < grp.N) ? expB[pr.b_exp + gn] : 0;
        double h = (double)lev[jb][NL - 1][r];
#pragma unroll
        for (int L = NL - 2; L >= 0; --L) h = fma(h, kInvRadix, (double)lev[jb][L][r]);
        double v = ldexp(h * kInvLead2, ea + eb);
        if (ea == kNonFinite || eb == kNonFinite) v = __builtin_nan("");
        acc[jb][r] = fma(ga * pr.alpha, v, acc[jb][r]);
      }
  }

  const int gm = m0 + wave * 32 + l32;
  if (gm >= grp.M) return;
#pragma unroll
  for (int jb = 0; jb < 2; ++jb)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int gn = n0 + jb * 32 + (r & 3) + 8 * (r >> 2) + 4 * lh;
      if (gn >= grp.N) continue;
      const int64_t o = grp.off_c + gm + (int64_t)gn * grp.ldc;
      double out = acc[jb][r];
      if (beta != 0.0) out = fma(beta, Cg[o], out);
      Cg[o] = out;
    }
}

}  // namespace

int split_slices(int prec) { return prec == 2 ? kSliceCount : 0; }

hipError_t launch_slice(hipStream_t s, int prec, int cplx, int nsets, const SliceBlock* blocks,
                        const int2* wgmap, int64_t grid, const SliceOperands& ops) {
  if (grid == 0) return hipSuccess;
  if (prec != 2) return hipErrorNotSupported;
  const dim3 g((unsigned)grid), blk(kThreads);
  if (cplx) {
    ITG_LAUNCH((slice_max_kernel<1>), g, blk, 0, s, blocks, wgmap, ops);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    ITG_LAUNCH((slice_kernel<kSliceCount, 1>), g, blk, 0, s, blocks, wgmap, ops, nsets);
  } else {
    ITG_LAUNCH((slice_max_kernel<0>), g, blk, 0, s, blocks, wgmap, ops);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    ITG_LAUNCH((slice_kernel<kSliceCount, 0>), g, blk, 0, s, blocks, wgmap, ops, 1);
  }
  return hipGetLastError();
}

int64_t slice_block_wgs(int32_t rows, int32_t kparts) {
  return (int64_t)((rows + kSliceRows - 1) / kSliceRows) * kparts;
}

namespace {
template <int CPLX, int WM>
hipError_t launch_sg(hipStream_t s, int comp, int sets, const SplitPair* pairs,
                     const GemmGroup* groups, const int2* tiles, int64_t grid,
                     const void* planesA, const void* planesB, const void* expA,
                     const void* expB, void* C, double ga, double beta, int aim) {
  const int ns = CPLX ? 3 : 1;
  ITG_LAUNCH((split_gemm_kernel<kSliceCount, CPLX, WM>), dim3((unsigned)grid),
                     dim3(128 * WM), 0, s, pairs, groups, tiles, (const int8_t*)planesA,
                     (const int8_t*)planesB, (const int32_t*)expA, (const int32_t*)expB,
                     (double*)C, ga, beta, CPLX ? comp : 0, CPLX ? sets : 0, ns, ns, aim);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_split_gemm(hipStream_t s, int prec, int cplx, int wmb, int comp, int sets,
                             const SplitPair* pairs, const GemmGroup* groups, const int2* tiles,
                             int64_t grid, const void* planesA, const void* planesB,
                             const void* expA, const void* expB, void* C, double ga,
                             double beta, int aim) {
  if (grid == 0) return hipSuccess;
  if (prec != 2) return hipErrorNotSupported;
  if (wmb == 3) {  // A in registers (real data; 64 x 64 tile table)
    if (cplx || aim) return hipErrorInvalidValue;
    ITG_LAUNCH((split_gemm_areg_kernel<kSliceCount>), dim3((unsigned)grid), dim3(128), 0,
                       s, pairs, groups, tiles, (const int8_t*)planesA, (const int8_t*)planesB,
                       (const int32_t*)expA, (const int32_t*)expB, (double*)C, ga, beta);
    return hipGetLastError();
  }
  if (wmb != 1 && wmb != 2) return hipErrorInvalidValue;
  if (cplx)
    return wmb == 2 ? launch_sg<1, 4>(s, comp, sets, pairs, groups, tiles, grid, planesA, planesB,
                                      expA, expB, C, ga, beta, aim)
                    : launch_sg<1, 2>(s, comp, sets, pairs, groups, tiles, grid, planesA, planesB,
                                      expA, expB, C, ga, beta, aim);
  return wmb == 2 ? launch_sg<0, 4>(s, comp, sets, pairs, groups, tiles, grid, planesA, planesB,
                                    expA, expB, C, ga, beta, aim)
                  : launch_sg<0, 2>(s, comp, sets, pairs, groups, tiles, grid, planesA, planesB,
                                    expA, expB, C, ga, beta, aim);
}

}  // namespace itg
