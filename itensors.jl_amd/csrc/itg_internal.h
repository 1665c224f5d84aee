// Internal declarations shared by the host planner and the HIP kernels of
// libitg.so. Nothing here is part of the C ABI (include/itg.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <tuple>
#include <utility>
#include <string>
#include <vector>

namespace itg {

constexpr int kMaxModes = 8;  // modes per tensor after validation
constexpr int kMaxKModes = 8;

// ------------------------------------------------------------ descriptors
// Column-major tensors; offsets and strides are in ELEMENTS of the dtype.

// One block of a grouped permutation: dst = alpha*perm(src) + beta*dst,
// described in the DST block's mode order (dst mode 0 fastest) after mode
// merging (nm modes, uniform per launch): the dst dims with their division
// constants and the src stride of every dst mode, all computed on the host,
// so the kernels walk the modes with compile-time indices only.
struct PermBlock {
  int64_t off_src;            // < 0: no source block (its values are zero)
  int64_t off_dst;
  int64_t sstr[kMaxModes];    // src stride of dst mode i
  uint32_t ddim[kMaxModes];   // dst dims
  uint32_t dmul[kMaxModes];   // x / ddim[i] == (umulhi(x, dmul[i]) + x) >> dshr[i], x < 2^31
  uint32_t dshr[kMaxModes];
  int32_t nelem;
  int32_t fresh;              // 1: the dst block is new (written, never read)
};
// Division constants of d >= 1 for dividends below 2^31 (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994,
// fig. 4.1): q = (umulhi(x, mul) + x) >> shr.
inline void fastdiv_consts(uint32_t d, uint32_t* mul, uint32_t* shr) {
  uint32_t l = 0;
  while (l < 32 && (uint64_t(1) << l) < d) ++l;
  *shr = l;
  *mul = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
}

// Small-K direct contraction: every output element of a group (output block)
// sums, over the group's pairs in plan order, alpha_p * sum_k A[..]*B[..].
struct DirectPair {
  int64_t off_a;
  int64_t off_b;
  double alpha_re, alpha_im;
  int32_t cstr_a[kMaxModes];  // stride in A of each R mode (0 if from B)
  int32_t cstr_b[kMaxModes];  // stride in B of each R mode (0 if from A)
  int32_t kofs;               // start of this pair's (ka, kb) offset list
  int32_t nk;                 // number of contracted elements
};
struct DirectGroup {
  int64_t off_c;
  int64_t wg0;     // first workgroup of the group
  int32_t cdim[kMaxModes];
  int32_t nelem;
  int32_t p0, p1;  // pair range
  int32_t pad;
};

// Grouped GEMM: C (M x N, column-major, ldc = M) of a group is
// sum over its pairs p of alpha_p * op(A_p) * op(B_p).
struct GemmPair {
  int64_t off_a;
  int64_t off_b;
  int32_t lda, ldb;
  int32_t K;
  int32_t pad;
  double alpha_re, alpha_im;
};
struct GemmGroup {
  int64_t off_c;
  int32_t M, N;
  int32_t ldc;
  int32_t p0, p1;
  int32_t pad;
};

// ------------------------------------------------------------ kernel launch
// Events that the kernels launched by this thread record their own start
// and end into (set by the per-kernel timer around one timed section, null
// otherwise). The first launch of a section records `start` and clears it;
// every launch records `stop`, so the section spans from the first kernel's
// start to the last kernel's end, as the dispatches' own timestamps (no
// gaps between separately enqueued event packets).
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
  bool launched = false;
};
LaunchEvents& launch_events();

template <typename Tup, size_t... I>
inline void kernel_arg_ptrs(Tup& t, void** p, std::index_sequence<I...>) {
  ((p[I] = (void*)&std::get<I>(t)), ...);
}

// Every kernel of the library is launched through here. The thread's
// last-error state is cleared first: launchers report hipGetLastError()
// after the launch, and an error left behind by an unrelated earlier HIP
// call in the process (the host application's own) must not be reported as
// this launch's.
template <typename... KArgs, typename... Args>
inline hipError_t launch_kernel(void (*kernel)(KArgs...), dim3 grid, dim3 block, unsigned shmem,
                                hipStream_t s, Args... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  (void)hipGetLastError();
  std::tuple<KArgs...> vals(static_cast<KArgs>(args)...);
  void* ptrs[sizeof...(KArgs)];
  kernel_arg_ptrs(vals, ptrs, std::index_sequence_for<KArgs...>{});
  LaunchEvents& ev = launch_events();
  if (!ev.stop) return hipLaunchKernel((const void*)kernel, grid, block, ptrs, shmem, s);
  hipError_t e =
      hipExtLaunchKernel((const void*)kernel, grid, block, ptrs, shmem, s, ev.start, ev.stop, 0);
  ev.start = nullptr;
  ev.launched = true;
  return e;
}
#define ITG_LAUNCH(kernel, grid, block, shmem, stream, ...) \
  (void)::itg::launch_kernel(kernel, dim3(grid), dim3(block), shmem, stream, __VA_ARGS__)

// --------------------------------------------------------- kernel launchers
// All launchers enqueue on `stream` and return hipSuccess or the launch error.
// `esize` is 1 for Float64 and 2 for ComplexF64.

// Element types of the data-movement kernels.
enum EltCode { kEltF64 = 0, kEltC128 = 1, kEltF32 = 2, kEltC64 = 3 };
// Grouped permutation. perm[i] = src mode feeding dst mode i (0-based).
// tile_prefix has nblk+1 entries (workgroups before block b). elt: EltCode
// (0 / 1 = the cplx flag of the FP64 paths). nzflags (optional, nblk ints,
// cleared by the launcher): set to 1 for every block with a written value
// that is not zero.
hipError_t launch_permute(hipStream_t s, int elt, int nm, const int32_t* perm_host,
                          const PermBlock* blocks_dev, const int64_t* tile_prefix_dev,
                          int64_t nblk, int64_t total_tiles, int mode_kind,
                          const void* src, void* dst, double a_re, double a_im,
                          double b_re, double b_im, int32_t* nzflags = nullptr);
// Number of workgroups a block of the permutation needs (host helper that
// must agree with the kernel's tiling). ddim: the block's dst dims; q: the
// dst mode fed by src mode 0. mode_kind: 0 = src mode 0 stays dst mode 0
// (streamed), 1 = 2-D tile transpose of dst mode 0 and dst mode q.
int64_t permute_block_tiles(int nm, int q, const uint32_t* ddim, int mode_kind);

// One block of a grouped strided copy (element offsets and strides; modes
// in traversal order, mode 0 fastest; nelem = product of dims < 2^31).
struct CopyBlock {
  int64_t off_src, off_dst;
  int64_t nelem;
  int32_t dims[kMaxModes];
  int64_t sstr[kMaxModes], dstr[kMaxModes];
};
hipError_t launch_copy_blocks(hipStream_t s, int cplx, int nm, const CopyBlock* blocks_dev,
                              const int64_t* prefix_dev, int64_t nblk, int64_t total_wgs,
                              const void* src, void* dst);
int64_t copy_block_wgs(int64_t nelem);  // workgroups of one block (host helper)

// wg_group: the group of every workgroup (total_tiles entries).
hipError_t launch_direct(hipStream_t s, int cplx, int nr, const DirectPair* pairs,
                         const DirectGroup* groups, const int32_t* koffs,
                         const int32_t* wg_group, int64_t total_tiles, const void* A,
                         const void* B, void* C, unsigned flags, double ga_re, double ga_im,
                         double beta_re, double beta_im);
constexpr int kDirectElemsPerWG = 256 * 4;

// Fiber-stationary direct contraction (small contracted extent, one operand
// much smaller than the other). "A" is the large operand, "B" the small one;
// a fiber is one value of A's free (uncontracted) multi-index. A group is a
// set of A blocks and output blocks that share their A-free sector: every
// output element of the group is sum_p alpha_p sum_k A_p[fiber, k] B_p[k, g]
// over the group's pairs p of that output block, g running over B's free
// multi-index. One thread owns a fiber and writes all of its outputs, so A
// is fetched from HBM once.
constexpr int kFibersPerWG = 256 * 4;
struct FiberGroup {
  int64_t nfib;              // fibers (product of fdim)
  int64_t wg0;               // first workgroup of the group
  int32_t fdim[kMaxModes];   // extents of A's free modes, in A's mode order
  int32_t r0, r1;            // FiberOut range
};
struct FiberOut {            // one output block of a group
  int64_t off_c;
  int32_t cstr[kMaxModes];   // output strides of the A-free modes (fdim order)
  int32_t g0, ng;            // gofs_c[g0 .. g0 + ng): output offsets of the B-free elements
  int32_t p0, p1;            // FiberPair range (plan order)
};
struct FiberPair {
  int64_t off_a, off_b;
  double alpha_re, alpha_im;
  int32_t astr[kMaxModes];   // A strides of the A-free modes (fdim order)
  int32_t kofs, nk;          // koffs[2*(kofs+k)], [..+1]: A / B offsets of contracted index k
  int32_t gb0;               // gofs_b[gb0 + g]: B offset of B-free element g
  int32_t pad;
};
// Matrix form of a group (when it has at most kFiberSlots distinct A values
// per fiber and kFiberOuts outputs per fiber): slot s = one (A block,
// contracted index) of the group, output o = one (output block, B-free
// index); out_o = sum_s a_s W[s][o] with W[s][o] = alpha_p B_p[k(s), g(o)]
// for the pair p linking them (0 if none). Each workgroup builds its
// group's W in LDS, loads all slots of its fibers at once (A read exactly
// once) and writes every output.
constexpr int kFiberSlots = 16, kFiberOuts = 32;
constexpr int kFibersPerWGMV = 256 * 8;  // 8 rounds of one fiber per thread
struct FiberMVGroup {
  int64_t nfib, wg0;
  int32_t fdim[kMaxModes];
  int32_t nslot, nout;
  int32_t s0, o0;            // FiberSlot / FiberOutElem ranges
  int32_t w0, nw;            // FiberCoef range
};
struct FiberSlot {           // A offset = off_a + sum_i c_i * astr[i] + ka
  int64_t off_a;
  int32_t astr[kMaxModes];
  int32_t ka, pad;
};
struct FiberOutElem {        // C offset = off_c + sum_i c_i * cstr[i] + go
  int64_t off_c;
  int32_t cstr[kMaxModes];
  int32_t go;
  uint32_t mask;             // bit s: slot s feeds this output (others are not read,
                             // so a non-finite A value reaches only its own outputs)
};
struct FiberCoef {           // W[slot][out] = alpha * B[off_b]
  int64_t off_b;
  double alpha_re, alpha_im;
  int32_t slot, out;
};
// nf: number of A-free modes (1..kMaxModes). flags bit 0 / 1: conj(A) / conj(B).
// smax: slot bound the kernel is unrolled for (4, 8 or 16; >= every group's nslot).
hipError_t launch_fiber_mv(hipStream_t s, int cplx, int nf, int smax, const FiberMVGroup* groups,
                           const FiberSlot* slots, const FiberOutElem* outs,
                           const FiberCoef* coefs, const int32_t* wg_group, int64_t total_wgs,
                           const void* A, const void* B, void* C, unsigned flags, double ga_re,
                           double ga_im, double beta_re, double beta_im);
// Two consecutive matrix-form fiber contractions fused, Y = X*B1 then
// Z = Y*B2, without storing Y. A super-fiber is one value of the modes that
// are fiber modes of both steps (the other fiber modes of either step have
// unit sector sizes, so a super-fiber is one fiber of each step). Per
// super-group: x = the X value of every step-1 slot (step-1 groups in
// order), y = every Y value a step-2 slot reads (in step-2 slot order), z =
// every step-2 output; y_v = sum_u x_u W1[v][u] over linked u in order,
// z_q = sum_v y_v W2[q][v] over linked v in order: the same operations in
// the same order as the two separate passes, so Z is bit-identical. Slot
// and output strides are given over the shared modes (fdim order).
constexpr int kChainX = 16, kChainY = 16, kChainZ = 32;
constexpr int kFibersPerWGChain = 256 * 4;  // 4 rounds: measured 3 % faster than 8 or 12
struct ChainGroup {
  int64_t nfib, wg0;
  int32_t fdim[kMaxModes];     // extents of the shared modes (1 past nf)
  int32_t nx, ny, nz;
  int32_t x0, y0, z0;          // FiberSlot (x) / ymask (y) / FiberOutElem (z) ranges
  int32_t w10, nw1, w20, nw2;  // coefs1 (slot = x, out = y) / coefs2 (slot = y, out = z)
  int32_t pad;
};
// xm, ym: unroll bounds (8 or 16). B1 / B2: the small operands of the steps.
hipError_t launch_fiber_chain(hipStream_t s, int cplx, int nf, int xm, int ym,
                              const ChainGroup* groups, const FiberSlot* slots,
                              const FiberOutElem* zouts, const uint32_t* ymask,
                              const FiberCoef* coefs1, const FiberCoef* coefs2,
                              const int32_t* wg_group, int64_t total_wgs, const void* X,
                              const void* B1, const void* B2, void* Z);

// Block-sparse tensor times a block-diagonal one (DiagBlockSparse: blocks
// (k, ..., k) storing only their diagonals), grouped over the output
// blocks: a workgroup writes kDiagChunk consecutive elements of one output
// block. For output element r and pair p (T1 block, D block):
//   u0 >= 0 (D has open modes): j = r's coordinate on R mode u0; the term
//     is d_j * T1[off1 + sum_i c_i str1[i] + j * kstr] when every R mode in
//     umask also has coordinate j and j < dlen, else 0;
//   u0 < 0 (D fully contracted): the sum over j < dlen of the same term.
// Pairs of one output block are summed in plan order; R = alpha * sum.
constexpr int kDiagChunk = 1024;
struct DiagOut {
  int64_t offR, nelem, wg0;
  int32_t dims[kMaxModes];
  int32_t p0, p1;            // DiagPair range
};
struct DiagPair {
  int64_t off1, offD, kstr;
  int32_t str1[kMaxModes];   // T1 stride of each R mode taken from T1 (0 for D's)
  int32_t u0, dlen;
  uint32_t umask;
  int32_t pad;
};
// nr: R modes (1..kMaxModes; an order-0 output is one mode of size 1).
// uniform: D is one value (u_re, u_im) on every diagonal element; D unused.
hipError_t launch_diag(hipStream_t s, int cplx, int nr, const DiagOut* outs,
                       const DiagPair* pairs, const int32_t* wg_out, int64_t total_wgs,
                       const void* T1, const void* D, int uniform, double u_re, double u_im,
                       double a_re, double a_im, void* R);

// nf: number of A-free modes (1..kMaxModes). flags bit 0 / 1: conj(A) / conj(B).
hipError_t launch_fiber(hipStream_t s, int cplx, int nf, const FiberGroup* groups,
                        const FiberOut* outs, const FiberPair* pairs, const int32_t* koffs,
                        const int32_t* gofs_c, const int32_t* gofs_b, const int32_t* wg_group,
                        int64_t total_wgs, const void* A, const void* B, void* C, unsigned flags,
                        double ga_re, double ga_im, double beta_re, double beta_im);
// Workgroups a direct-contraction group needs (kDirectElemsPerWG output
// elements each; host helper that must agree with the kernel).
int64_t direct_group_tiles(const DirectGroup* g, int nr);

// Grouped GEMM (FP64 MFMA). cfg selects the tile shape; ta/tb: A stored
// K x M / B stored N x K. `tiles` holds one {group, tile index in group} per
// workgroup ({-1, -1} = idle slot), laid out so that workgroup w runs on
// XCD w % 8.
struct GemmTileCfg {
  int bm, bn;
};
int gemm_num_cfgs();
GemmTileCfg gemm_cfg(int cfg);
bool gemm_cfg_ok(int cfg, int cplx);
hipError_t launch_gemm_f64(hipStream_t s, int cfg, int cplx, int ta, int tb,
                           const GemmPair* pairs, const GemmGroup* groups, const int2* tiles,
                           int64_t grid_size, const void* A, const void* B, void* C,
                           unsigned flags, double ga_re, double ga_im, double beta_re,
                           double beta_im);

// Split-K form of the grouped FP64 GEMM: the GEMM kernel writes S partial
// sums of each output block (one per slice of its contracted extent) into a
// workspace, and this kernel adds them in slice order and writes C = sum +
// beta * C (beta == 0 does not read C). One entry per output block: partials
// at P[poff + s*M*N], s = 0..S-1, each an M x N column-major block; the
// block's elements are e0 .. e0 + M*N - 1 of the launch.
struct KsReduce {
  int64_t poff, off_c, e0;
  int32_t M, N, ldc, S;
};
// Dot-like split-K plans (every output block has at most kKsDotMN elements):
// the partial blocks come from a streaming vector-ALU kernel, one workgroup
// per entry of `groups` (the split tables), instead of the MFMA GEMM.
constexpr int kKsDotMN = 16;
// one: every block has a single element (inner products).
hipError_t launch_ksplit_dot(hipStream_t s, int cplx, int one, int ta, int tb,
                             const GemmPair* pairs, const GemmGroup* groups, int64_t ngroups,
                             const void* A, const void* B, void* P, unsigned flags,
                             double ga_re, double ga_im);
// Skinny split-K plans (every output block has a side of at most kKsSkinny:
// matrix times a few vectors): rows = (entry of `groups`, first row of the
// long side) per workgroup, kKsSkinnyRows[kc] rows each. kc: the long side's
// rows are contiguous along the contracted values.
constexpr int kKsSkinny = 4;
constexpr int kKsSkinnyRows[2] = {256, 16};
hipError_t launch_ksplit_skinny(hipStream_t s, int cplx, int kc, int ta, int tb,
                                const GemmPair* pairs, const GemmGroup* groups, const int2* rows,
                                int64_t nrows, const void* A, const void* B, void* P,
                                unsigned flags, double ga_re, double ga_im);
// max_slices: the largest S of the entries (many slices: a wave per element).
hipError_t launch_ksplit_reduce(hipStream_t s, int cplx, const KsReduce* red, int nred,
                                int64_t total, int max_slices, const void* P, void* C,
                                double beta_re, double beta_im);

// ------------------------------------------------ split-precision GEMM
// FP64 operands are split into kSliceCount int8 digits per element (radix
// kSliceRadix, the first digit scaled by kSliceLead / 2^E with one E per row:
// M row of A, N column of B); see kernels_split.hip.
constexpr int kSliceModes = 4;  // row / K modes of a sliced operand after merging
constexpr int kSliceCount = 6;        // int8 digits per FP64 value
constexpr double kSliceRadix = 254.0;  // digits in [-127, 127]
constexpr double kSliceLead = 127.0;   // |first digit| <= 127
// Longest contracted extent of one pair of the split GEMM: the level sums
// (at most kSliceCount products of two digits per k) stay below 2^31 in
// INT32 for 2*K (the complex form stacks two K halves): 6 * 127^2 * 16384 <
// 2^31. Longer pairs are cut into K chunks summed in FP64.
constexpr int kSplitKChunk = 8192;

// One block of an operand to slice: rows x K read through mode strides.
constexpr int kSliceRows = 64;    // rows per slicing workgroup
constexpr int kSliceKPart = 128;  // K elements per slicing workgroup
struct SliceBlock {
  int64_t off;        // element offset of the block in the FP64 operand
  int64_t plane_off;  // byte offset of slice plane 0 (R x Kpad bytes per plane,
                      // chunk-major: 16-byte K chunk c of row r at (c*R + r)*16)
  int64_t exp_off;    // index of the block's first row exponent
  int64_t pmax_off;   // first partial exponent (kparts x R ints, part-major)
  int32_t R, K, Kpad;
  int32_t kparts;     // ceil(Kpad / kSliceKPart): workgroups along K
  int32_t nrm, nkm;   // number of row / K modes
  int32_t side;       // operand: 0 = A, 1 = B (one launch slices both)
  int32_t pad;
  int32_t rdim[kSliceModes], rstr[kSliceModes];
  int32_t kdim[kSliceModes], kstr[kSliceModes];
};
struct SplitPair {
  int64_t a_planes, b_planes;    // byte offsets of slice plane 0 of the block (per slice set)
  int64_t a_koff, b_koff;        // byte offset of this K chunk inside each plane (k0 * rows)
  int64_t a_pstride, b_pstride;  // bytes between consecutive slice planes
  int32_t a_exp, b_exp;          // first row exponent of the A block / B block
  int32_t kpad, pad;
  double alpha, alpha_im;        // per-pair factor (the kernel takes one part per launch)
};

int split_slices(int prec);  // slices per value for a precision mode (0: none)
// Workgroups of one block: ceil(rows / kSliceRows) x kparts.
int64_t slice_block_wgs(int32_t rows, int32_t kparts);
// The buffers of the two operands of a split GEMM; SliceBlock::side picks.
struct SliceOperands {
  const double* src[2];  // FP64 operand (complex: interleaved)
  int8_t* planes[2];     // slice planes
  int32_t* exps[2];      // row exponents
  int32_t* pmax[2];      // partial row exponents
};
// Two launches over the same grid (the blocks of both operands, A's first):
// partial row maxima (as exponents, into pmax), then the slices (planes) and
// the row exponents (exps). Complex operands: nsets = 3 slice sets (re, im,
// -im); plane buffers hold nsets x the per-set bytes.
// Workgroup w slices tile wgmap[w] = (block, tile index within the block:
// row tile * kparts + K part).
hipError_t launch_slice(hipStream_t s, int prec, int cplx, int nsets, const SliceBlock* blocks,
                        const int2* wgmap, int64_t grid, const SliceOperands& ops);
// Tile shape of the split GEMM: 64 x 64 (wmb = 1, 256 threads, tile table of
// gemm cfg 0) or 128 x 64 (wmb = 2, 512 threads, tile table of gemm cfg 2).
// Complex: one launch per output component comp (0 = re, 1 = im); `sets`
// packs the slice sets (0 re, 1 im, 2 -im) of A and B for the two K halves:
// bits 0-1 A half 0, 2-3 B half 0, 4-5 A half 1, 6-7 B half 1. Both complex
// operands are sliced with 3 sets.
// aim: take each pair's factor from alpha_im instead of alpha (the second
// pass of a product with complex per-pair factors).
hipError_t launch_split_gemm(hipStream_t s, int prec, int cplx, int wmb, int comp, int sets,
                             const SplitPair* pairs, const GemmGroup* groups, const int2* tiles,
                             int64_t grid, const void* planesA, const void* planesB,
                             const void* expA, const void* expB, void* C, double ga,
                             double beta, int aim = 0);
// x[i] = beta * x[i] for i < n (complex when cplx).
hipError_t launch_scale(hipStream_t s, int cplx, void* x, int64_t n, double b_re, double b_im);
// Float32 / ComplexF32 at the boundary: x64[i] = x32[i] (n elements; cplx:
// interleaved pairs), and y32[i] = round(x64[i] + beta * y32[i]) with the sum
// formed in FP64 (beta == 0 does not read y32).
hipError_t launch_widen(hipStream_t s, int cplx, int64_t n, const void* x32, void* x64);
hipError_t launch_narrow(hipStream_t s, int cplx, int64_t n, const void* x64, void* y32,
                         double b_re, double b_im);
// y = x converted from EltCode `from` to the wider EltCode `to` (n elements;
// hipErrorInvalidValue for a pair that is not a widening).
hipError_t launch_convert(hipStream_t s, int from, int to, int64_t n, const void* x, void* y);
// Complex y[i] = alpha * (x[i] + i * x2[i]) + beta * y[i] for i < n (x2 may
// be null: taken as 0; beta == 0 does not read y).
hipError_t launch_axpby_join(hipStream_t s, int64_t n, double a_re, double a_im, const void* x,
                             const void* x2, double b_re, double b_im, void* y);

}  // namespace itg
