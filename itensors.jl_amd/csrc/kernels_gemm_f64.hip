// Grouped FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// One launch covers every output block ("group") of a contraction plan: a
// dense contraction is one group with one pair; a QN block-sparse step has
// one group per output block and one pair per allowed (block1, block2).
// A workgroup owns one BM x BN tile of one group and walks the group's pairs
// in plan order, accumulating in registers, then writes the tile once:
//
//   C_g = beta*C_g (dense entry only) + sum_{p in g} alpha_p * op(A_p) op(B_p)
//
// so every output element has exactly one writer and a fixed summation
// order (deterministic, no atomics).
//
// Scheduling: the host builds a tile table in which workgroup w runs on XCD
// w % 8 (the hardware's round-robin dispatch). Each XCD receives whole groups
// (or large chunks of a big group) so the A and B panels of a group are
// fetched once into that XCD's L2 and shared by its tiles; chunks are spread
// over the XCDs largest-first to balance the work.
//
// Tiles: 256 threads = 4 wave64s in a 2x2 grid; each wave owns a
// (BM/2) x (BN/2) sub-tile as (BM/32) x (BN/32) MFMA 16x16 blocks of FP64
// accumulators. A and B K-slices (BK = 16) are double-buffered in LDS: the
// next slice is fetched into registers while the current one is multiplied,
// then written to the other buffer, with one barrier per slice. The MFMA is
// issued as D = B^T A^T so that each lane's accumulators run down a column
// of C, which makes the column-major epilogue stores contiguous.
//
// Complex data (interleaved re/im) uses the 4-real-product form on the same
// tiles, with conj(A)/conj(B) folded into the sign of the imaginary parts.
// Replaces BLAS dgemm/zgemm behind mul!! (NDTensors/src/abstractarray/mul.jl:1-12,
// NDTensors/src/array/mul.jl:1-4) and the per-pair loop of
// NDTensors/src/blocksparse/contract_generic.jl:78-129.
#include <algorithm>

#include "itg_internal.h"

namespace itg {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int BK = 16;
constexpr int PAD = 1;

template <int BM, int BN, int CPLX, int TA, int TB>
__global__ void __launch_bounds__(kThreads)
    gemm_f64_kernel(const GemmPair* __restrict__ pairs, const GemmGroup* __restrict__ groups,
                    const int2* __restrict__ tiles, const double* __restrict__ Ag,
                    const double* __restrict__ Bg, double* __restrict__ Cg, unsigned flags,
                    double ga_re, double ga_im, double beta_re, double beta_im) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int NA = BM * BK / kThreads;  // A elements loaded per thread
  constexpr int NB = BN * BK / kThreads;
  constexpr int NPLANE = CPLX ? 2 : 1;

  __shared__ double As[2][NPLANE][BK][BM + PAD];
  __shared__ double Bs[2][NPLANE][BK][BN + PAD];

  const int2 tt = tiles[blockIdx.x];
  if (tt.x < 0) return;  // padding slot of the XCD-balanced table
  const GemmGroup grp = groups[tt.x];
  const int tiles_m = (grp.M + BM - 1) / BM;
  const int tm = tt.y % tiles_m, tn = tt.y / tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;
  const int l16 = lane & 15, lk = lane >> 4;

  const double sa = (flags & 1u) ? -1.0 : 1.0;  // conj(A): negate Im(A)
  const double sb = (flags & 2u) ? -1.0 : 1.0;

  v4d acc[NPLANE][FM][FN];
#pragma unroll
  for (int c = 0; c < NPLANE; ++c)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[c][i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  double ra[NPLANE][NA], rb[NPLANE][NB];

  for (int p = grp.p0; p < grp.p1; ++p) {
    const GemmPair pr = pairs[p];
    const int K = pr.K;
    const int M = grp.M, N = grp.N;
    const double al_re = ga_re * pr.alpha_re - ga_im * pr.alpha_im;
    const double al_im = ga_re * pr.alpha_im + ga_im * pr.alpha_re;

    // Loads are unconditional from clamped (in-block) addresses so that they
    // issue back to back and their latency overlaps the MFMAs; out-of-range
    // elements are zeroed, and alpha / conj applied, when the registers are
    // written to LDS.
    unsigned ina = 0u, inb = 0u;
    auto load_regs = [&](int k0) {
      ina = inb = 0u;
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const int idx = tid + r * kThreads;
        const int m = TA ? idx / BK : idx % BM;
        const int k = TA ? idx % BK : idx / BM;
        const int gm = m0 + m, gk = k0 + k;
        if (gm < M && gk < K) ina |= 1u << r;
        const int gmc = min(gm, M - 1), gkc = min(gk, K - 1);
        const int64_t off = TA ? ((int64_t)gmc * pr.lda + gkc) : ((int64_t)gkc * pr.lda + gmc);
        if (CPLX) {
          const double2 v = reinterpret_cast<const double2*>(Ag)[pr.off_a + off];
          ra[0][r] = v.x;
          ra[NPLANE - 1][r] = v.y;
        } else {
          ra[0][r] = Ag[pr.off_a + off];
        }
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int idx = tid + r * kThreads;
        const int n = TB ? idx % BN : idx / BK;
        const int k = TB ? idx / BN : idx % BK;
        const int gn = n0 + n, gk = k0 + k;
        if (gn < N && gk < K) inb |= 1u << r;
        const int gnc = min(gn, N - 1), gkc = min(gk, K - 1);
        const int64_t off = TB ? ((int64_t)gkc * pr.ldb + gnc) : ((int64_t)gnc * pr.ldb + gkc);
        if (CPLX) {
          const double2 v = reinterpret_cast<const double2*>(Bg)[pr.off_b + off];
          rb[0][r] = v.x;
          rb[NPLANE - 1][r] = v.y;
        } else {
          rb[0][r] = Bg[pr.off_b + off];
        }
      }
    };
    auto store_lds = [&](int buf) {
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const int idx = tid + r * kThreads;
        const int m = TA ? idx / BK : idx % BM;
        const int k = TA ? idx % BK : idx / BM;
        const bool in = (ina >> r) & 1u;
        if (CPLX) {
          const double x = in ? ra[0][r] : 0.0;
          const double y = in ? sa * ra[NPLANE - 1][r] : 0.0;
          As[buf][0][k][m] = al_re * x - al_im * y;
          As[buf][NPLANE - 1][k][m] = al_re * y + al_im * x;
        } else {
          As[buf][0][k][m] = in ? al_re * ra[0][r] : 0.0;
        }
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int idx = tid + r * kThreads;
        const int n = TB ? idx % BN : idx / BK;
        const int k = TB ? idx / BN : idx % BK;
        const bool in = (inb >> r) & 1u;
        Bs[buf][0][k][n] = in ? rb[0][r] : 0.0;
        if (CPLX) Bs[buf][NPLANE - 1][k][n] = in ? sb * rb[NPLANE - 1][r] : 0.0;
      }
    };

    if (K <= 0) continue;
    // every wave is past the previous pair's last read of buffer 0 (the
    // barrier that closes each K-slice), so it can be refilled here
    load_regs(0);
    store_lds(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += BK) {
      const bool more = k0 + BK < K;
      if (more) load_regs(k0 + BK);
      // every fragment of the slice is read before the first MFMA (each kk
      // into its own registers), so the reads of later kk land while the
      // MFMAs of earlier ones run; left to itself the scheduler reuses one
      // register set and waits for each kk's reads with lgkmcnt(0)
      constexpr int NKK = BK / 4;
      double af[NKK][NPLANE][FM], bf[NKK][NPLANE][FN];
#pragma unroll
      for (int q = 0; q < NKK; ++q)
#pragma unroll
        for (int c = 0; c < NPLANE; ++c) {
#pragma unroll
          for (int i = 0; i < FM; ++i)
            af[q][c][i] = As[buf][c][4 * q + lk][wm0 + i * 16 + l16];
#pragma unroll
          for (int j = 0; j < FN; ++j)
            bf[q][c][j] = Bs[buf][c][4 * q + lk][wn0 + j * 16 + l16];
        }
#pragma unroll
      for (int q = 0; q < NKK; ++q)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            if (CPLX) {
              // re += Bre*Are - Bim*Aim ; im += Bre*Aim + Bim*Are
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[q][0][j], af[q][0][i],
                                                                  acc[0][i][j], 0, 0, 0);
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                  -bf[q][NPLANE - 1][j], af[q][NPLANE - 1][i], acc[0][i][j], 0, 0, 0);
              acc[NPLANE - 1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                  bf[q][0][j], af[q][NPLANE - 1][i], acc[NPLANE - 1][i][j], 0, 0, 0);
              acc[NPLANE - 1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                  bf[q][NPLANE - 1][j], af[q][0][i], acc[NPLANE - 1][i][j], 0, 0, 0);
            } else {
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[q][0][j], af[q][0][i],
                                                                  acc[0][i][j], 0, 0, 0);
            }
          }
      __builtin_amdgcn_sched_group_barrier(0x100, NKK * NPLANE * (FM + FN), 0);  // DS reads
      __builtin_amdgcn_sched_group_barrier(0x008, NKK * FM * FN * (CPLX ? 4 : 1), 0);  // MFMAs
      __builtin_amdgcn_sched_barrier(0);
      // the other buffer was last read in the previous slice, before the
      // barrier that ended it
      if (more) store_lds(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }

  // Epilogue. v_mfma_f64_16x16x4_f64 leaves D[i][j] in lane j + 16*(i%4),
  // register i/4 (rows interleaved across the four lane groups), so with
  // D = (A B)^T this lane holds C[m = l16][n = lk + 4*r] of each 16x16 block.
  const bool use_beta = (beta_re != 0.0 || beta_im != 0.0);
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int gm = m0 + wm0 + i * 16 + l16;
    if (gm >= grp.M) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gn = n0 + wn0 + j * 16 + lk + 4 * r;
        if (gn >= grp.N) continue;
        const int64_t o = grp.off_c + gm + (int64_t)gn * grp.ldc;
        if (CPLX) {
          double2 out = make_double2(acc[0][i][j][r], acc[NPLANE - 1][i][j][r]);
          if (use_beta) {
            const double2 c = reinterpret_cast<double2*>(Cg)[o];
            out.x += beta_re * c.x - beta_im * c.y;
            out.y += beta_re * c.y + beta_im * c.x;
          }
          reinterpret_cast<double2*>(Cg)[o] = out;
        } else {
          double out = acc[0][i][j][r];
          if (use_beta) out = fma(beta_re, Cg[o], out);
          Cg[o] = out;
        }
      }
    }
  }
}

// The block of output element e, found by binary search over the blocks'
// first elements (a split-K plan has few blocks).
__device__ inline int ksplit_block(const KsReduce* red, int nred, int64_t e) {
  int lo = 0, hi = nred - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (red[mid].e0 <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// WIDE = false: one thread per output element, adding the partials in slice
// order. WIDE = true (many slices, few elements: inner products, small
// cotangents): one wave per output element, lane l adding slices l, l + 64,
// ... in order, then a butterfly over the lanes -- a fixed order for a given
// slice count, so the result is deterministic. Beta is applied as the
// GEMM's epilogue does.
template <int CPLX, bool WIDE>
__global__ void __launch_bounds__(kThreads)
    ksplit_reduce_kernel(const KsReduce* __restrict__ red, int nred, int64_t total,
                         const double* __restrict__ P, double* __restrict__ C, double beta_re,
                         double beta_im) {
  const bool use_beta = (beta_re != 0.0 || beta_im != 0.0);
  constexpr int kPer = WIDE ? 64 : 1;  // threads per element
  const int lane = WIDE ? (int)(threadIdx.x & 63) : 0;
  const int64_t stride = (int64_t)gridDim.x * (kThreads / kPer);
  for (int64_t e = (int64_t)blockIdx.x * (kThreads / kPer) + threadIdx.x / kPer; e < total;
       e += stride) {
    const KsReduce r = red[ksplit_block(red, nred, e)];
    const int64_t loc = e - r.e0, mn = (int64_t)r.M * r.N;
    double vr = 0.0, vi = 0.0;
    // unrolled so that the loads of several slices are in flight at once
#pragma unroll 8
    for (int q = lane; q < r.S; q += kPer) {
      if (CPLX) {
        const double2 w = reinterpret_cast<const double2*>(P)[r.poff + q * mn + loc];
        vr = q == lane ? w.x : vr + w.x;
        vi = q == lane ? w.y : vi + w.y;
      } else {
        const double w = P[r.poff + q * mn + loc];
        vr = q == lane ? w : vr + w;
      }
    }
    if (WIDE) {
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        vr += __shfl_xor(vr, m);
        if (CPLX) vi += __shfl_xor(vi, m);
      }
      if (lane != 0) continue;
    }
    const int64_t n = loc / r.M, m = loc - n * r.M;
    const int64_t o = r.off_c + m + n * r.ldc;
    if (CPLX) {
      double2 v = make_double2(vr, vi);
      if (use_beta) {
        const double2 c = reinterpret_cast<double2*>(C)[o];
        v.x += beta_re * c.x - beta_im * c.y;
        v.y += beta_re * c.y + beta_im * c.x;
      }
      reinterpret_cast<double2*>(C)[o] = v;
    } else {
      double v = vr;
      if (use_beta) v = fma(beta_re, C[o], v);
      C[o] = v;
    }
  }
}

// Split-K partial blocks of dot-like plans (every output block has at most
// kKsDotMN elements: inner products, norms): one workgroup per (block,
// slice) entry of the split tables streams its sub-pairs' operands once on
// the vector ALUs (a 64 x 64 MFMA tile would compute 4096 products per
// useful one), each thread summing every kThreads-th contracted value into
// M x N accumulators (MNC = 1: one output, four values in flight per thread;
// MNC = kKsDotMN: up to that many outputs); the workgroup then adds its
// threads' sums (a fixed tree) and writes the partial block. Per pair, the
// sum is scaled by that pair's factor times the global one, as the GEMM
// scales its A operand.
template <int CPLX, int MNC>
__global__ void __launch_bounds__(kThreads)
    ksplit_dot_kernel(const GemmPair* __restrict__ pairs, const GemmGroup* __restrict__ groups,
                      const double* __restrict__ Ag, const double* __restrict__ Bg,
                      double* __restrict__ P, int ta, int tb, unsigned flags, double ga_re,
                      double ga_im) {
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int U = MNC == 1 ? 4 : 1;  // contracted values per thread in flight
  __shared__ double red[kThreads / 64][MNC * NP];
  const GemmGroup grp = groups[blockIdx.x];
  const int M = grp.M, MN = MNC == 1 ? 1 : grp.M * grp.N;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double sa = (flags & 1u) ? -1.0 : 1.0;  // conj(A): negate Im(A)
  const double sb = (flags & 2u) ? -1.0 : 1.0;
  double tot[MNC][NP];
#pragma unroll
  for (int i = 0; i < MNC; ++i)
#pragma unroll
    for (int c = 0; c < NP; ++c) tot[i][c] = 0.0;
  for (int p = grp.p0; p < grp.p1; ++p) {
    const GemmPair pr = pairs[p];
    double acc[MNC][NP];
#pragma unroll
    for (int i = 0; i < MNC; ++i)
#pragma unroll
      for (int c = 0; c < NP; ++c) acc[i][c] = 0.0;
    for (int k0 = tid; k0 < pr.K; k0 += U * kThreads) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u * kThreads;
        if (U > 1 && k >= pr.K) break;
#pragma unroll
        for (int i = 0; i < MNC; ++i) {
          if (MNC > 1 && i >= MN) continue;
          const int m = i % M, n = i / M;
          const int64_t oa =
              pr.off_a + (ta ? (int64_t)m * pr.lda + k : (int64_t)k * pr.lda + m);
          const int64_t ob =
              pr.off_b + (tb ? (int64_t)k * pr.ldb + n : (int64_t)n * pr.ldb + k);
          if (CPLX) {
            const double2 a = reinterpret_cast<const double2*>(Ag)[oa];
            const double2 b = reinterpret_cast<const double2*>(Bg)[ob];
            const double ai = sa * a.y, bi = sb * b.y;
            acc[i][0] = fma(a.x, b.x, fma(-ai, bi, acc[i][0]));
            acc[i][NP - 1] = fma(a.x, bi, fma(ai, b.x, acc[i][NP - 1]));
          } else {
            acc[i][0] = fma(Ag[oa], Bg[ob], acc[i][0]);
          }
        }
      }
    }
    const double al_re = ga_re * pr.alpha_re - ga_im * pr.alpha_im;
    const double al_im = ga_re * pr.alpha_im + ga_im * pr.alpha_re;
#pragma unroll
    for (int i = 0; i < MNC; ++i) {
      if (CPLX) {
        tot[i][0] += al_re * acc[i][0] - al_im * acc[i][NP - 1];
        tot[i][NP - 1] += al_re * acc[i][NP - 1] + al_im * acc[i][0];
      } else {
        tot[i][0] += al_re * acc[i][0];
      }
    }
  }
  // threads -> waves (butterfly), waves -> one thread per output (wave order)
#pragma unroll
  for (int i = 0; i < MNC; ++i)
#pragma unroll
    for (int c = 0; c < NP; ++c) {
      double v = tot[i][c];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
      if (lane == 0) red[wave][i * NP + c] = v;
    }
  __syncthreads();
  if (tid < MN) {
    double v[NP];
#pragma unroll
    for (int c = 0; c < NP; ++c) {
      v[c] = red[0][tid * NP + c];
      for (int w = 1; w < kThreads / 64; ++w) v[c] += red[w][tid * NP + c];
    }
    const int m = tid % M, n = tid / M;
    const int64_t o = grp.off_c + m + (int64_t)n * grp.ldc;
    if (CPLX)
      reinterpret_cast<double2*>(P)[o] = make_double2(v[0], v[NP - 1]);
    else
      P[o] = v[0];
  }
}

// Split-K partial blocks of skinny plans (every output block has one side of
// at most kKsSkinny: matrix times vector or a few vectors). Workgroup w
// takes consecutive rows of the long side of entry rows[w].x of the split
// tables, starting at rows[w].y, and writes those rows of the partial
// block. KC = 0 (the long side's rows run along the long side in memory):
// 256 rows, a thread owns one and walks the entry's contracted values in
// order, four in flight, so a wave's loads are 64 consecutive rows. KC = 1
// (each row is contiguous along the contracted values): 16 rows, a wave
// takes four of them with its lanes striding along the contracted values,
// then adds each row's lane sums with a butterfly. The short side's values
// are the same for every row (they come from the caches). Per pair the sums
// are scaled by that pair's factor times the global one.
template <int CPLX, int KC>
__global__ void __launch_bounds__(kThreads)
    ksplit_skinny_kernel(const GemmPair* __restrict__ pairs, const GemmGroup* __restrict__ groups,
                         const int2* __restrict__ rows, const double* __restrict__ Ag,
                         const double* __restrict__ Bg, double* __restrict__ P, int ta, int tb,
                         unsigned flags, double ga_re, double ga_im) {
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int CS = kKsSkinny;
  constexpr int NR = KC ? 4 : 1;  // rows per thread (KC = 1: per wave)
  const int2 t = rows[blockIdx.x];
  const GemmGroup grp = groups[t.x];
  const bool mrows = grp.N <= CS;  // the long side is M (A's rows), else N
  const int R = mrows ? grp.M : grp.N, C = mrows ? grp.N : grp.M;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = KC ? t.y + wave * NR : t.y + (int)threadIdx.x;
  if (r0 >= R) return;  // KC = 1: the whole wave (r0 is the same on every lane)
  const double sa = (flags & 1u) ? -1.0 : 1.0;  // conj(A): negate Im(A)
  const double sb = (flags & 2u) ? -1.0 : 1.0;
  double tot[NR][CS][NP];
#pragma unroll
  for (int j = 0; j < NR; ++j)
#pragma unroll
    for (int c = 0; c < CS; ++c)
#pragma unroll
      for (int q = 0; q < NP; ++q) tot[j][c][q] = 0.0;
  for (int p = grp.p0; p < grp.p1; ++p) {
    const GemmPair pr = pairs[p];
    // element (m, k) of A and (n, k) of B, as the GEMM kernel indexes them
    auto aoff = [&](int m, int k) {
      return pr.off_a + (ta ? (int64_t)m * pr.lda + k : (int64_t)k * pr.lda + m);
    };
    auto boff = [&](int n, int k) {
      return pr.off_b + (tb ? (int64_t)k * pr.ldb + n : (int64_t)n * pr.ldb + k);
    };
    double acc[NR][CS][NP];
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
      for (int c = 0; c < CS; ++c)
#pragma unroll
        for (int q = 0; q < NP; ++q) acc[j][c][q] = 0.0;
    // KC = 0: k = 0, 1, ... four at a time; KC = 1: k = lane, lane + 64, ...
    constexpr int U = KC ? 1 : 4;
    for (int k0 = KC ? lane : 0; k0 < pr.K; k0 += KC ? 64 : U) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u;
        if (!KC && k >= pr.K) break;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          const int rr = min(r0 + j, R - 1);  // rows past R: computed, not stored
          const int64_t ox = mrows ? aoff(rr, k) : boff(rr, k);
          if (CPLX) {
            double2 x = reinterpret_cast<const double2*>(mrows ? Ag : Bg)[ox];
            x.y *= mrows ? sa : sb;
#pragma unroll
            for (int c = 0; c < CS; ++c) {
              if (c >= C) break;
              double2 y = reinterpret_cast<const double2*>(mrows ? Bg : Ag)[mrows ? boff(c, k)
                                                                                  : aoff(c, k)];
              y.y *= mrows ? sb : sa;
              acc[j][c][0] = fma(x.x, y.x, fma(-x.y, y.y, acc[j][c][0]));
              acc[j][c][NP - 1] = fma(x.x, y.y, fma(x.y, y.x, acc[j][c][NP - 1]));
            }
          } else {
            const double x = (mrows ? Ag : Bg)[ox];
#pragma unroll
            for (int c = 0; c < CS; ++c) {
              if (c >= C) break;
              const double y = (mrows ? Bg : Ag)[mrows ? boff(c, k) : aoff(c, k)];
              acc[j][c][0] = fma(x, y, acc[j][c][0]);
            }
          }
        }
      }
    }
    const double al_re = ga_re * pr.alpha_re - ga_im * pr.alpha_im;
    const double al_im = ga_re * pr.alpha_im + ga_im * pr.alpha_re;
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
      for (int c = 0; c < CS; ++c) {
        double ar = acc[j][c][0], ai = acc[j][c][NP - 1];
        if (KC) {  // the row's sum over the wave's lanes (a fixed tree)
#pragma unroll
          for (int m = 32; m >= 1; m >>= 1) {
            ar += __shfl_xor(ar, m);
            if (CPLX) ai += __shfl_xor(ai, m);
          }
        }
        if (CPLX) {
          tot[j][c][0] += al_re * ar - al_im * ai;
          tot[j][c][NP - 1] += al_re * ai + al_im * ar;
        } else {
          tot[j][c][0] += al_re * ar;
        }
      }
  }
  if (KC && lane != 0) return;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (r0 + j >= R) break;
#pragma unroll
    for (int c = 0; c < CS; ++c) {
      if (c >= C) break;
      const int m = mrows ? r0 + j : c, n = mrows ? c : r0 + j;
      const int64_t o = grp.off_c + m + (int64_t)n * grp.ldc;
      if (CPLX)
        reinterpret_cast<double2*>(P)[o] = make_double2(tot[j][c][0], tot[j][c][NP - 1]);
      else
        P[o] = tot[j][c][0];
    }
  }
}

struct CfgEntry {
  int bm, bn;
};
constexpr CfgEntry kCfgs[] = {{64, 64}, {128, 128}, {128, 64}, {64, 128}};

template <int BM, int BN, int CPLX>
hipError_t launch_tt(hipStream_t s, int ta, int tb, dim3 grid, const GemmPair* pairs,
                     const GemmGroup* groups, const int2* tiles, const double* A,
                     const double* B, double* C, unsigned flags, double ar, double ai, double br,
                     double bi) {
  dim3 block(kThreads);
  if (!ta && !tb)
    ITG_LAUNCH((gemm_f64_kernel<BM, BN, CPLX, 0, 0>), grid, block, 0, s, pairs, groups,
                       tiles, A, B, C, flags, ar, ai, br, bi);
  else if (!ta && tb)
    ITG_LAUNCH((gemm_f64_kernel<BM, BN, CPLX, 0, 1>), grid, block, 0, s, pairs, groups,
                       tiles, A, B, C, flags, ar, ai, br, bi);
  else if (ta && !tb)
    ITG_LAUNCH((gemm_f64_kernel<BM, BN, CPLX, 1, 0>), grid, block, 0, s, pairs, groups,
                       tiles, A, B, C, flags, ar, ai, br, bi);
  else
    ITG_LAUNCH((gemm_f64_kernel<BM, BN, CPLX, 1, 1>), grid, block, 0, s, pairs, groups,
                       tiles, A, B, C, flags, ar, ai, br, bi);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_ksplit_reduce(hipStream_t s, int cplx, const KsReduce* red, int nred,
                                int64_t total, int max_slices, const void* P, void* C,
                                double beta_re, double beta_im) {
  if (total <= 0) return hipSuccess;
  if (nred <= 0) return hipErrorInvalidValue;
  const bool wide = max_slices >= 64;
  const int64_t per_block = wide ? kThreads / 64 : kThreads;
  const int64_t blocks = std::min<int64_t>((total + per_block - 1) / per_block, 4096);
#define ITG_KS_LAUNCH(c, w)                                                                  \
  ITG_LAUNCH((ksplit_reduce_kernel<c, w>), dim3((unsigned)blocks), dim3(kThreads), 0, s, red, \
             nred, total, (const double*)P, (double*)C, beta_re, beta_im)
  if (cplx && wide)
    ITG_KS_LAUNCH(1, true);
  else if (cplx)
    ITG_KS_LAUNCH(1, false);
  else if (wide)
    ITG_KS_LAUNCH(0, true);
  else
    ITG_KS_LAUNCH(0, false);
#undef ITG_KS_LAUNCH
  return hipGetLastError();
}

hipError_t launch_ksplit_dot(hipStream_t s, int cplx, int one, int ta, int tb,
                             const GemmPair* pairs, const GemmGroup* groups, int64_t ngroups,
                             const void* A, const void* B, void* P, unsigned flags,
                             double ga_re, double ga_im) {
  if (ngroups <= 0) return hipSuccess;
#define ITG_DOT_LAUNCH(c, mn)                                                                  \
  ITG_LAUNCH((ksplit_dot_kernel<c, mn>), dim3((unsigned)ngroups), dim3(kThreads), 0, s, pairs, \
             groups, (const double*)A, (const double*)B, (double*)P, ta, tb, flags, ga_re,     \
             ga_im)
  if (cplx && one)
    ITG_DOT_LAUNCH(1, 1);
  else if (cplx)
    ITG_DOT_LAUNCH(1, kKsDotMN);
  else if (one)
    ITG_DOT_LAUNCH(0, 1);
  else
    ITG_DOT_LAUNCH(0, kKsDotMN);
#undef ITG_DOT_LAUNCH
  return hipGetLastError();
}

hipError_t launch_ksplit_skinny(hipStream_t s, int cplx, int kc, int ta, int tb,
                                const GemmPair* pairs, const GemmGroup* groups, const int2* rows,
                                int64_t nrows, const void* A, const void* B, void* P,
                                unsigned flags, double ga_re, double ga_im) {
  if (nrows <= 0) return hipSuccess;
#define ITG_SK_LAUNCH(c, k)                                                                    \
  ITG_LAUNCH((ksplit_skinny_kernel<c, k>), dim3((unsigned)nrows), dim3(kThreads), 0, s, pairs, \
             groups, rows, (const double*)A, (const double*)B, (double*)P, ta, tb, flags,      \
             ga_re, ga_im)
  if (cplx && kc)
    ITG_SK_LAUNCH(1, 1);
  else if (cplx)
    ITG_SK_LAUNCH(1, 0);
  else if (kc)
    ITG_SK_LAUNCH(0, 1);
  else
    ITG_SK_LAUNCH(0, 0);
#undef ITG_SK_LAUNCH
  return hipGetLastError();
}

int gemm_num_cfgs() { return (int)(sizeof(kCfgs) / sizeof(kCfgs[0])); }
GemmTileCfg gemm_cfg(int cfg) { return GemmTileCfg{kCfgs[cfg].bm, kCfgs[cfg].bn}; }
bool gemm_cfg_ok(int cfg, int cplx) { return !cplx || cfg == 0 || cfg == 2; }

hipError_t launch_gemm_f64(hipStream_t s, int cfg, int cplx, int ta, int tb,
                           const GemmPair* pairs, const GemmGroup* groups, const int2* tiles,
                           int64_t grid_size, const void* A, const void* B, void* C,
                           unsigned flags, double ga_re, double ga_im, double beta_re,
                           double beta_im) {
  if (grid_size == 0) return hipSuccess;
  if (!gemm_cfg_ok(cfg, cplx)) return hipErrorInvalidValue;
  dim3 grid((unsigned)grid_size);
  const double* a = (const double*)A;
  const double* b = (const double*)B;
  double* c = (double*)C;
  if (cplx) {
    // complex accumulators double the register tile: keep tiles <= 128x64
    if (cfg == 2)
      return launch_tt<128, 64, 1>(s, ta, tb, grid, pairs, groups, tiles, a, b, c, flags, ga_re,
                                   ga_im, beta_re, beta_im);
    return launch_tt<64, 64, 1>(s, ta, tb, grid, pairs, groups, tiles, a, b, c, flags, ga_re,
                                ga_im, beta_re, beta_im);
  }
  switch (cfg) {
    case 1:
      return launch_tt<128, 128, 0>(s, ta, tb, grid, pairs, groups, tiles, a, b, c, flags, ga_re,
                                    ga_im, beta_re, beta_im);
    case 2:
      return launch_tt<128, 64, 0>(s, ta, tb, grid, pairs, groups, tiles, a, b, c, flags, ga_re,
                                   ga_im, beta_re, beta_im);
    case 3:
      return launch_tt<64, 128, 0>(s, ta, tb, grid, pairs, groups, tiles, a, b, c, flags, ga_re,
                                   ga_im, beta_re, beta_im);
    default:
      return launch_tt<64, 64, 0>(s, ta, tb, grid, pairs, groups, tiles, a, b, c, flags, ga_re,
                                  ga_im, beta_re, beta_im);
  }
}

}  // namespace itg
