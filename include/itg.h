/*
 * itg.h -- C ABI of libitg.so, the MI355X (gfx950) backend for NDTensors'
 * tensor contraction (dense and QN block-sparse).
 *
 * Conventions shared by every entry point
 *  - Tensors are column-major (first mode fastest), exactly as NDTensors'
 *    `Dense` / `BlockSparse` storages hold them. Complex elements are
 *    interleaved (re, im) pairs of doubles, i.e. Julia's ComplexF64.
 *  - All data pointers are DEVICE pointers owned by the caller. The library
 *    owns only handles, plans, device-side plan tables and its workspace;
 *    each is released by the matching destroy call.
 *  - Labels follow the ncon convention that
 *    ITensors.compute_contraction_labels produces (src/indexset.jl:672-707):
 *    a label shared by T1 and T2 is contracted (ITensors makes it negative),
 *    every other label appears exactly once in R. Label values are Julia Int.
 *  - Scalars (alpha, beta) are passed as two doubles {re, im}; the imaginary
 *    part is ignored for real dtypes.
 *  - beta == 0 never reads R: R may hold NaN/Inf on entry
 *    (NDTensors/test/test_dense.jl:237-259).
 *  - Every function returns an itg_status; no C++ exception crosses the ABI.
 *    itg_last_error() gives the message of the last failure on that handle.
 *  - A handle is not thread-safe: use one handle per host thread, or
 *    serialise calls. Work is enqueued on the handle's stream (default:
 *    the null stream); calls return once the work is enqueued unless noted.
 */
#ifndef ITG_H
#define ITG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int itg_status;
enum {
  ITG_OK = 0,
  ITG_ERR_INVALID = 1,     /* bad argument (shapes, labels, pointers)       */
  ITG_ERR_UNSUPPORTED = 2, /* valid for NDTensors but not for this backend: */
                           /* the caller falls back to its own path, as the */
                           /* cuTENSOR extension does on CUTENSORError      */
                           /* (NDTensors/ext/NDTensorscuTENSORExt/contract.jl:45-53) */
  ITG_ERR_HIP = 3,         /* a HIP runtime call failed                     */
  ITG_ERR_NOMEM = 4        /* device allocation of workspace failed         */
};

/* Element types (Julia eltypes). Float32 / ComplexF32 are accepted by
 * itg_permute, itg_bs_permute(_ex), itg_contract_dense(_ex), itg_bs_execute
 * and itg_bs_diag_execute: permutations move them as they are; contractions
 * run in FP64 (MFMA) on widened copies of the operands, whatever the
 * precision mode, and round once when R is written (R = round(alpha*T1*T2 +
 * beta*R), the sum formed in FP64). The other entry points take FP64 data
 * and return ITG_ERR_UNSUPPORTED for these. */
typedef enum {
  ITG_F32 = 0,  /* Float32    */
  ITG_F64 = 1,  /* Float64    */
  ITG_C64 = 2,  /* ComplexF32 */
  ITG_C128 = 3  /* ComplexF64 */
} itg_dtype;

/* Arithmetic form of the GEMM-shaped work. */
typedef enum {
  ITG_PREC_FP64 = 0,       /* FP64 MFMA (v_mfma_f64_16x16x4_f64), FP64 accumulate */
  ITG_PREC_SPLIT_BF16 = 1, /* reserved for BF16 slices; returns                 */
                           /* ITG_ERR_UNSUPPORTED in this version                 */
  ITG_PREC_SPLIT_INT8 = 2  /* FP64 operands split into 6 int8 digits of radix 254 */
                           /* (|digit| <= 127) under a scale 2^E / 127 per row    */
                           /* (A) / per column (B); the 21 digit products of      */
                           /* weight >= 254^-5 run on INT8 MFMA with exact INT32  */
                           /* sums per weight level, recombined and summed in     */
                           /* FP64 (about 1e-13 relative Frobenius error)         */
} itg_prec;

/* Execution flags for itg_bs_execute / itg_contract_dense_ex. */
enum {
  ITG_FLAG_CONJ_A = 1u, /* use conj(T1) (dag of complex data, src/itensor.jl:1806-1816) */
  ITG_FLAG_CONJ_B = 2u, /* use conj(T2)                                                  */
  /* Fixed operands in split precision (itg_bs_execute / _noalloc only;
   * ignored in FP64 and by the other entry points). No NDTensors
   * counterpart: an eigensolver applies one effective Hamiltonian many times
   * with the same environment tensors, whose digit planes need not be
   * recomputed each time.
   *   KEEP_SLICES_1 / _2: slice T1 / T2 into storage owned by the plan (and
   *     its per-pair factor set) and keep it, remembering the data pointer;
   *   REUSE_SLICES_1 / _2: do not slice T1 / T2; use the planes a KEEP call
   *     of this plan left for the same pointer and element type
   *     (ITG_ERR_INVALID if there are none). The caller guarantees the data
   *     has not changed since that KEEP call. */
  ITG_FLAG_KEEP_SLICES_1 = 4u,
  ITG_FLAG_KEEP_SLICES_2 = 8u,
  ITG_FLAG_REUSE_SLICES_1 = 16u,
  ITG_FLAG_REUSE_SLICES_2 = 32u
};

typedef struct itg_handle itg_handle;
typedef struct itg_bs_plan itg_bs_plan;

/* ---------------------------------------------------------------- handle */

/* Create a handle bound to HIP device `device`. No reference counterpart:
 * it holds what the Julia extension would keep per task (stream, workspace). */
itg_status itg_init(int device, itg_handle** out);
itg_status itg_destroy(itg_handle* h);
const char* itg_last_error(const itg_handle* h);
/* Enqueue subsequent work on `stream` (a hipStream_t; NULL = null stream). */
itg_status itg_set_stream(itg_handle* h, void* stream);
itg_status itg_synchronize(itg_handle* h);
/* Library version, "major.minor.patch". */
const char* itg_version(void);

/* Tile shape of the split-precision GEMM for later contractions on this
 * handle: 0 = per plan (the autotuned shape, else chosen by size),
 * 1 = 64 x 64, 2 = 128 x 64, 3 = 64 x 64 with the A operand held in
 * registers (real data; complex plans use 1). All shapes give the same
 * result bit for bit. No counterpart in NDTensors, whose BLAS call has no
 * tile shape; for tests and tuning. ITG_ERR_INVALID outside 0..3. */
itg_status itg_set_split_tile(itg_handle* h, int tile);

/* Kernel family of the contractions planned on this handle from now on (a
 * block-sparse plan keeps the family it was created with; dense calls are
 * planned per family). The counterpart of
 * NDTensors' gemm_backend switch (NDTensors/src/abstractarray/tensoralgebra/
 * contract.jl:15-24: :Auto, :BLAS, :Generic):
 *   0 = Auto    -- by shape: grouped GEMM on MFMA (TTGT) when the
 *                  flop-weighted mean contracted extent is >= 24, else the
 *                  fiber or strided kernels;
 *   1 = BLAS    -- the grouped GEMM whenever a mode is contracted;
 *   2 = Generic -- the generic strided kernel for every contraction (each
 *                  output element summed from strided reads: no
 *                  permutation, no MFMA).
 * Results agree to rounding across the three. ITG_ERR_INVALID otherwise. */
itg_status itg_set_gemm_backend(itg_handle* h, int backend);

/* Block-pair order of the block-sparse plans created on this handle from now
 * on: the counterpart of NDTensors' threaded block-sparse switch, which picks
 * the algorithm of contract_blockoffsets (NDTensors/src/blocksparse/
 * contract.jl:42-55).
 *   ITG_BS_ORDER_SEQUENTIAL -- T1's blocks outer, T2's inner
 *                              (contract_sequential.jl:26-39);
 *   ITG_BS_ORDER_THREADED   -- contract_blocks(::Algorithm"threaded_threads")
 *                              (contract_threaded.jl:2-73): the same when T1
 *                              has more blocks than T2, else T2's blocks
 *                              outer and T1's inner.
 * The order fixes the output's BlockOffsets (blocks placed in order of first
 * appearance, contract_generic.jl:23-30) and the order of the sums into each
 * output block. The reference takes the threaded algorithm when
 * using_threaded_blocksparse() && Threads.nthreads() > 1; a Julia caller
 * passes that condition. DiagBlockSparse plans have one order (the
 * reference's diagonal contraction has no threaded form). ITG_ERR_INVALID
 * for another value. */
typedef enum {
  ITG_BS_ORDER_SEQUENTIAL = 0,
  ITG_BS_ORDER_THREADED = 1
} itg_bs_order;
itg_status itg_set_blocksparse_order(itg_handle* h, int order);

/* Per-kernel timing: the device counterpart of NDTensors' global
 * `timer::TimerOutput` (NDTensors/src/NDTensors.jl:96-99) and its
 * `@timeit_debug` sections around BLAS.gemm! and the permutations
 * (NDTensors/src/abstractarray/tensoralgebra/contract.jl:61-63,125-184).
 * While enabled, every kernel launch the handle makes outside stream capture
 * is bracketed by HIP events on the handle's stream. itg_timer_enable(h, 1)
 * clears the totals and starts; (h, 0) stops. itg_timer_read synchronizes
 * the stream and reports entry i (0 <= i < itg_timer_count) as a kernel
 * name, its launch count and its summed duration in milliseconds; `name`
 * stays valid until the next timer call on the handle. */
itg_status itg_timer_enable(itg_handle* h, int enable);
int64_t itg_timer_count(itg_handle* h);
itg_status itg_timer_read(itg_handle* h, int64_t i, const char** name, int64_t* calls,
                          double* total_ms);

/* ------------------------------------------------------------ HIP graphs */
/* Record the handle's subsequent launches into a graph instead of running
 * them, then replay the whole sequence with one launch: the captured form of
 * an effective-Hamiltonian apply that an eigensolver repeats. A handle on
 * the null stream records on a stream of its own (after waiting for the null
 * stream); otherwise the handle's stream is captured. Every call recorded
 * must have run once on this handle before (its plan tables, per-pair
 * factor tables and workspace then exist: building them is not
 * capturable, and a call that would needs makes itg_graph_end fail). The
 * graph reads the caller's buffers at the addresses it was recorded with,
 * and the plans it uses must outlive it; the library keeps every workspace
 * and per-pair factor table a graph may reference until the handle / plan is
 * destroyed. While a capture is open, calls on the handle only record.
 * No NDTensors counterpart (its contractions run eagerly); the replacement
 * for the hipGraph of the north-star batching pattern at the boundary. */
typedef struct itg_graph itg_graph;
itg_status itg_graph_begin(itg_handle* h);
itg_status itg_graph_end(itg_handle* h, itg_graph** out);
/* Launch the graph on the handle's stream (stream-ordered, returns at once). */
itg_status itg_graph_launch(itg_handle* h, itg_graph* g);
itg_status itg_graph_destroy(itg_graph* g);

/* ---------------------------------------------------------- device memory */
/* Device buffers for callers without their own GPU array package (the Julia
 * extension's ItgVector storage type, which stands where CuArray/ROCArray
 * stand for the reference's GPU extensions, NDTensors/ext/NDTensorsAMDGPUExt/
 * adapt.jl:9-13). Copies are stream-ordered; the host-side ones return after
 * the data has arrived. */
itg_status itg_malloc(itg_handle* h, size_t bytes, void** ptr);
itg_status itg_free(itg_handle* h, void* ptr);
itg_status itg_memcpy_htod(itg_handle* h, void* dst, const void* src, size_t bytes);
itg_status itg_memcpy_dtoh(itg_handle* h, void* dst, const void* src, size_t bytes);
itg_status itg_memcpy_dtod(itg_handle* h, void* dst, const void* src, size_t bytes);
/* dst[i] = src[i] for i < n, converted from element type `from` to the wider
 * type `to` (Float32 -> Float64 / ComplexF32 / ComplexF64, Float64 ->
 * ComplexF64, ComplexF32 -> ComplexF64; the same type copies). The device
 * side of the promotion by allocation the dense contract! does for mixed
 * element types (NDTensors/src/dense/tensoralgebra/contract.jl:196-211,
 * NDTensors/ext/NDTensorscuTENSORExt/contract.jl:34-41). ITG_ERR_INVALID
 * for a narrowing pair. Stream-ordered. */
itg_status itg_convert(itg_handle* h, itg_dtype from, const void* src, itg_dtype to, void* dst,
                       int64_t n);

/* --------------------------------------------------------------- permute */

/* dst = alpha * permutedims(src, perm) + beta * dst.
 * `dims` are the dimensions of src; perm is 1-based (Julia) and
 * dst's mode i is src's mode perm[i].
 * Replaces NDTensors.permutedims!(::Exposed{<:Array}, ...) with @strided
 * (NDTensors/src/array/permutedims.jl:5-24) and the alpha/beta variants used
 * by the scalar contraction paths
 * (NDTensors/src/abstractarray/tensoralgebra/contract.jl:88-113). */
itg_status itg_permute(itg_handle* h, itg_dtype dtype, int ndim,
                       const int64_t* dims, const int32_t* perm,
                       const void* src, void* dst,
                       const double* alpha, const double* beta);

/* ------------------------------------------------------ dense contraction */

/* R = alpha * contract(T1, labelsT1, T2, labelsT2) + beta * R, with R's mode
 * order given by labelsR.
 * Replaces NDTensors.contract!(R::DenseTensor, labelsR, T1::DenseTensor,
 * labelsT1, T2::DenseTensor, labelsT2, alpha, beta)
 * (NDTensors/src/dense/tensoralgebra/contract.jl:160-216), i.e. the method
 * the cuTENSOR extension overrides (NDTensors/ext/NDTensorscuTENSORExt/
 * contract.jl:15-24). Covers the outer-product (N1+N2==NR) and
 * scalar (nnz==1) cases of that method too. All three tensors must have the
 * same dtype (the Julia side promotes first, as
 * dense/tensoralgebra/contract.jl:196-211 does). */
itg_status itg_contract_dense(itg_handle* h, itg_dtype dtype,
                              int NR, const int64_t* dimsR, const int64_t* labelsR, void* R,
                              int N1, const int64_t* dims1, const int64_t* labels1, const void* T1,
                              int N2, const int64_t* dims2, const int64_t* labels2, const void* T2,
                              const double* alpha, const double* beta, itg_prec prec);

/* Same as itg_contract_dense with ITG_FLAG_CONJ_* flags (used by the adjoint,
 * ext/ITensorsChainRulesCoreExt/itensor.jl:94-103, where dag(x) conjugates). */
itg_status itg_contract_dense_ex(itg_handle* h, itg_dtype dtype,
                                 int NR, const int64_t* dimsR, const int64_t* labelsR, void* R,
                                 int N1, const int64_t* dims1, const int64_t* labels1, const void* T1,
                                 int N2, const int64_t* dims2, const int64_t* labels2, const void* T2,
                                 const double* alpha, const double* beta, itg_prec prec,
                                 unsigned flags);

/* -------------------------------------------------- block-sparse contract */

/* Build the contraction plan of two BlockSparse tensors.
 * Replaces contraction_output + contract_blockoffsets(Algorithm"sequential")
 * (NDTensors/src/blocksparse/contract.jl:20-55,
 *  NDTensors/src/blocksparse/contract_sequential.jl:1-41).
 *
 *  blocks1  : N1 x nblk1 block coordinates, 1-based, block b at
 *             blocks1[b*N1 .. b*N1+N1-1] (the UInt tuples of Block{N1}),
 *             in the insertion order of T1's BlockOffsets.
 *  offs1    : nblk1 0-based element offsets of those blocks in T1's data.
 *  nsec1    : N1 numbers of sectors per mode; secdims1 concatenates the
 *             sector sizes of every mode (BlockDims, blockdims.jl:8-19).
 *  (same for T2)
 *  labels1/labels2/labelsR as for itg_contract_dense.
 *
 * The output blocks come out in the reference's order: first appearance while
 * iterating block1 over T1's blocks (outer) and block2 over T2's blocks
 * (inner); offsets accumulate in that order. Sector sizes of a contracted
 * mode must agree between T1 and T2. */
itg_status itg_bs_plan_create(itg_handle* h,
                              int N1, int64_t nblk1, const int64_t* blocks1, const int64_t* offs1,
                              const int64_t* nsec1, const int64_t* secdims1, const int64_t* labels1,
                              int N2, int64_t nblk2, const int64_t* blocks2, const int64_t* offs2,
                              const int64_t* nsec2, const int64_t* secdims2, const int64_t* labels2,
                              int NR, const int64_t* labelsR,
                              itg_bs_plan** out);

/* Output structure: number of blocks, total nnz, number of block pairs. */
itg_status itg_bs_plan_sizes(const itg_bs_plan* p, int64_t* nblkR, int64_t* nnzR,
                             int64_t* npairs);
/* blocksR: NR x nblkR 1-based coordinates; offsR: nblkR 0-based offsets. */
itg_status itg_bs_plan_blocks(const itg_bs_plan* p, int64_t* blocksR, int64_t* offsR);
/* pairs: 3 x npairs 0-based (block index in T1, in T2, in R), plan order. */
itg_status itg_bs_plan_pairs(const itg_bs_plan* p, int64_t* pairs);
/* How the plan runs (no reference counterpart; for callers choosing the
 * mode order of intermediates): info[0] = strategy (0 strided kernel, 1
 * fiber kernel, 2 grouped GEMM, -1 empty or diagonal plan); info[1..3] =
 * 1 when the GEMM permutes T1, T2, or its output before / after the GEMM
 * (FP64; split precision reads the operands through their strides). */
itg_status itg_bs_plan_info(const itg_bs_plan* p, int32_t* info);

/* R = alpha * (T1 * T2) on every output block of the plan (each output
 * block is overwritten by its first pair and accumulated by the rest, as
 * contract_generic.jl:78-129 does with beta = 0 then 1). R must have
 * nnzR elements. `pair_alpha` is NULL or npairs {re,im} per-pair factors
 * multiplying alpha (the slot compute_alpha fills for fermions,
 * contract_utilities.jl:2-15). An empty plan returns without work.
 * Replaces contract!(R, labelsR, T1, labelsT1, T2, labelsT2, plan)
 * (NDTensors/src/blocksparse/contract.jl:57-76) and its per-pair loop. */
itg_status itg_bs_execute(itg_handle* h, itg_bs_plan* p, itg_dtype dtype,
                          void* R, const void* T1, const void* T2,
                          const double* alpha, const double* pair_alpha,
                          itg_prec prec, unsigned flags);

itg_status itg_bs_plan_destroy(itg_bs_plan* p);

/* Size of the library's fixed device scratch buffer (64 MiB per GPU). */
#define ITG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* itg_bs_execute without any device allocation, for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc. No counterpart in the reference (NDTensors
 * allocates its outputs and temporaries as it goes). Same arguments and
 * results as itg_bs_execute, with three limits:
 *  - the workspace comes from a fixed device array of
 *    ITG_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 *    the process ends (its address is resolved per device by itg_init); a
 *    contraction needing more returns ITG_ERR_INVALID with the size;
 *  - the plan's device tables for this element type, precision and set of
 *    per-pair factors must exist: execute the plan once through
 *    itg_bs_execute first (ITG_ERR_INVALID otherwise);
 *  - Float64 / ComplexF64 only (ITG_ERR_UNSUPPORTED for Float32 /
 *    ComplexF32, whose widened copies would need memory).
 * Every call on one GPU shares that GPU's scratch array and it keeps the
 * last call's contents: calls that overlap on one GPU (several streams or
 * host threads) race, so serialise them. */
itg_status itg_bs_execute_noalloc(itg_handle* h, itg_bs_plan* p, itg_dtype dtype,
                                  void* R, const void* T1, const void* T2,
                                  const double* alpha, const double* pair_alpha,
                                  itg_prec prec, unsigned flags);

/* Block-sparse tensor T1 times a block-diagonal tensor D (DiagBlockSparse:
 * every stored block has equal block coordinates (k, ..., k) and holds only
 * its diagonal, min(block dims) elements at offsD[b]). Same arguments as
 * itg_bs_plan_create with D in the place of T2; the output structure
 * (itg_bs_plan_sizes / _blocks / _pairs) is the one the reference's
 * contraction_output gives. ITG_ERR_INVALID if a block of D is off the
 * block diagonal (the reference raises an error there too).
 * Replaces contraction_output + contract! for (BlockSparse, DiagBlockSparse)
 * in either order (NDTensors/src/blocksparse/diagblocksparse.jl:596-675),
 * reached from ITensor * ITensor with a delta / diag_itensor operand
 * (src/qn/qnitensor.jl:484-543). */
itg_status itg_bs_diag_plan_create(itg_handle* h,
                                   int N1, int64_t nblk1, const int64_t* blocks1,
                                   const int64_t* offs1, const int64_t* nsec1,
                                   const int64_t* secdims1, const int64_t* labels1,
                                   int ND, int64_t nblkD, const int64_t* blocksD,
                                   const int64_t* offsD, const int64_t* nsecD,
                                   const int64_t* secdimsD, const int64_t* labelsD,
                                   int NR, const int64_t* labelsR, itg_bs_plan** out);
/* R = alpha * (T1 * D) on every output block of a diagonal plan (blocks
 * overwritten; pairs of one output block summed in plan order). uniform:
 * NULL, or {re, im} of the one value on every diagonal element (delta,
 * UniformDiagBlockSparse; D is then unused). Per pair this is the dense
 * contract! of the T1 block with the diagonal block
 * (NDTensors/src/diag/tensoralgebra/contract.jl:103-160), without forming
 * the dense block. */
itg_status itg_bs_diag_execute(itg_handle* h, itg_bs_plan* p, itg_dtype dtype, void* R,
                               const void* T1, const void* D, const double* uniform,
                               const double* alpha);

/* Two consecutive block-sparse contractions fused into one pass over HBM:
 * Z = (X * B1) * B2 with p1 the plan of X * B1 (output Y) and p2 the plan
 * of Y * B2 (Y its first operand), without storing Y. Available when both
 * steps take the fiber path with few distinct operand values per fiber and
 * the open modes of Y that only one step iterates over have unit sector
 * sizes (the two MPO tensors of a two-site effective Hamiltonian applied
 * one after the other); otherwise create returns ITG_ERR_UNSUPPORTED and
 * the caller runs the two plans. Z holds p2's output structure. Results
 * are bit-identical to itg_bs_execute(p1) then itg_bs_execute(p2) with
 * alpha = 1, no per-pair factors and no conjugation, in every precision
 * mode (both steps compute in FP64). The chain holds its own tables; it
 * stays valid after p1 and p2 are destroyed.
 * Replaces two successive contract! calls on block-sparse tensors
 * (NDTensors/src/blocksparse/contract.jl:57-76) made by the left-to-right
 * fold of contract(::Vector{ITensor}) (src/tensor_operations/
 * tensor_algebra.jl:121-130, each step the A * B of tensor_algebra.jl:
 * 62-77), for the consecutive pairs that admit it. */
typedef struct itg_bs_chain itg_bs_chain;
itg_status itg_bs_chain_create(itg_handle* h, const itg_bs_plan* p1, const itg_bs_plan* p2,
                               itg_bs_chain** out);
itg_status itg_bs_chain_execute(itg_handle* h, itg_bs_chain* c, itg_dtype dtype, void* Z,
                                const void* X, const void* B1, const void* B2);
itg_status itg_bs_chain_destroy(itg_bs_chain* c);

/* Block-sparse permutedims: dst block b (same order and offsets as the caller
 * passes in offsR) = permutedims(src block b, perm). dims of each block are
 * N x nblk (block b at dims[b*N..]). Replaces permutedims!(R::BlockSparseTensor,
 * T::BlockSparseTensor, perm) (NDTensors/src/blocksparse/blocksparsetensor.jl:834-881)
 * for the common case of matching block lists. */
itg_status itg_bs_permute(itg_handle* h, itg_dtype dtype, int N, int64_t nblk,
                          const int64_t* blockdims, const int64_t* offs_src,
                          const int64_t* offs_dst, const int32_t* perm,
                          const void* src, void* dst);

/* The general block-sparse permutedims!(R, T, perm, f) with
 * f(r, t) = beta * r + alpha * t over the UNION of R's blocks and T's
 * permuted blocks (NDTensors/src/blocksparse/blocksparsetensor.jl:834-881):
 * the caller lists every block of the union (R's blocks in R's order, then
 * T's permuted blocks that R lacks, in T's order, as the reference iterates
 * `union(blocks_R, perm_blocks_T)`), with, per block b:
 *   blockdims[b*N ..]  the block's dims in T's mode order;
 *   offs_src[b]        its offset in T's data, or -1 when T has no such
 *                      block (t = 0: the block becomes beta * r);
 *   offs_dst[b]        its offset in R's data (R's storage already holds
 *                      room for the new blocks);
 *   fresh[b]           (fresh may be NULL = none) 1 when R had no such block:
 *                      r = 0 and the destination is written without being read.
 * nonzero (NULL or nblk ints) returns 1 for every block whose written values
 * are not all zero: the reference inserts a new block into R only when it is
 * nonzero, so the caller drops the fresh blocks flagged 0. beta == 0 never
 * reads R. Synchronous on the handle's stream. */
itg_status itg_bs_permute_ex(itg_handle* h, itg_dtype dtype, int N, int64_t nblk,
                             const int64_t* blockdims, const int64_t* offs_src,
                             const int64_t* offs_dst, const uint8_t* fresh,
                             const int32_t* perm, const void* src, void* dst,
                             const double* alpha, const double* beta, int32_t* nonzero);

/* Grouped strided block copy: for each block b (0 <= b < nblk) and every
 * multi-index x over dims[b*nmodes .. b*nmodes+nmodes),
 *   dst[dst_off[b] + sum_m x_m*dst_str[b*nmodes+m]]
 *       = src[src_off[b] + sum_m x_m*src_str[b*nmodes+m]]
 * with element offsets and strides (a ComplexF64 counts as one element).
 * src_len / dst_len are the lengths of the two device arrays in elements;
 * every access is checked against them on the host before anything runs
 * (ITG_ERR_INVALID otherwise). Destinations of different elements must not
 * coincide. This is the data movement of the block-sparse combiner
 * contraction: the per-block permutedims! of a block into its range of a
 * merged output block when combining (permutedims_combine,
 * NDTensors/src/blocksparse/blocksparsetensor.jl:571-630) and the copyto!
 * of a range of a merged block into an output block when uncombining
 * (uncombine, blocksparsetensor.jl:726-795). Synchronous on the handle's
 * stream; not for use under stream capture (it uploads its tables). */
itg_status itg_bs_copy(itg_handle* h, itg_dtype dtype, int64_t nblk, int nmodes,
                       const int64_t* dims, const int64_t* src_off, const int64_t* src_str,
                       const int64_t* dst_off, const int64_t* dst_str, const void* src,
                       int64_t src_len, void* dst, int64_t dst_len);

/* ---------------------------------------------------- batched contractions */

typedef struct itg_batch_plan itg_batch_plan;

/* nbatch independent dense contractions sharing one label pattern, each with
 * its own dimensions (dims1 is N1 x nbatch: item i at dims1[i*N1..]) and its
 * own caller-owned device buffers T1[i], T2[i], R[i] (host arrays of device
 * pointers, bound at creation). Each kernel of the contraction then runs as
 * one grouped launch over all items, which makes a batch of small networks
 * one captured launch per step instead of nbatch host round trips.
 * No reference counterpart: NDTensors contracts the items one by one (a host
 * loop over contract!, NDTensors/src/dense/tensoralgebra/contract.jl:160-216);
 * this is the batched form of itg_contract_dense for callers that hold many
 * small networks. */
itg_status itg_batch_plan_create(itg_handle* h, itg_dtype dtype, int64_t nbatch,
                                 int NR, const int64_t* dimsR, const int64_t* labelsR,
                                 void* const* R,
                                 int N1, const int64_t* dims1, const int64_t* labels1,
                                 const void* const* T1,
                                 int N2, const int64_t* dims2, const int64_t* labels2,
                                 const void* const* T2, itg_batch_plan** out);
/* R[i] = alpha * T1[i]*T2[i] for every item (beta = 0). */
itg_status itg_batch_execute(itg_handle* h, itg_batch_plan* p, const double* alpha,
                             itg_prec prec);
itg_status itg_batch_plan_destroy(itg_batch_plan* p);

/* ------------------------------------------------------------- autotuner */

/* Time the GEMM tile shapes on this plan's pairs and keep the fastest for
 * this plan and (dtype, prec); plans not tuned themselves use the choice
 * cached for their (shape bucket, dtype, prec). Synchronises the stream. */
itg_status itg_bs_autotune(itg_handle* h, itg_bs_plan* p, itg_dtype dtype,
                           void* R, const void* T1, const void* T2, itg_prec prec);

#ifdef __cplusplus
}
#endif

#endif /* ITG_H */
