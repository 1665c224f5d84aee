// Host-side contraction planner of libitg.so (no device code).
//
// A contraction (dense, or one QN block-sparse step) is lowered to a list of
// "pairs" (block of T1, block of T2 -> block of R) grouped by output block.
// The planner picks one of two strategies for the whole list:
//   DIRECT : small contracted extent; every output element is summed in place
//            from strided reads (no permutations).
//   FIBER  : DIRECT's shape class with one operand much larger than the
//            other: a thread owns a fiber of the large operand and writes
//            every output it feeds (the large operand crosses HBM once).
//   GEMM   : transpose-transpose-GEMM-transpose. Operands are used in place
//            when their modes are already grouped as (rows, contracted) or
//            (contracted, rows) -- the GEMM kernel takes either orientation --
//            and otherwise permuted once (grouped over blocks) into a
//            workspace. If R's modes interleave the two operands' free modes,
//            C is computed in (T1-free, T2-free) order and permuted into R.
// The choice of contracted-mode order and of C = T1*T2 versus C^T = T2^T*T1^T
// minimises the bytes permuted, in the spirit of NDTensors'
// ContractionProperties (NDTensors/src/tensoroperations/contraction_logic.jl:
// 283-651) but decided once per plan instead of once per block pair.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/itg.h"
#include "itg_internal.h"

namespace itg {

// Per-kernel HIP-event timing (enabled per handle). begin()/end() bracket one
// launch on stream s; fold() waits for the stream and adds the elapsed times
// to the per-name totals.
struct KernelTimer {
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<int64_t, double>> totals;  // calls, ms
  hipStream_t stream = nullptr;
  std::string open_name;
  hipEvent_t open_a = nullptr, open_b = nullptr;

  hipEvent_t take();
  void begin(const char* name, hipStream_t s);
  void end(hipStream_t s);
  hipError_t fold();
  void clear();
  ~KernelTimer();
};

struct Structure {
  int NR = 0, N1 = 0, N2 = 0;
  int r_src[kMaxModes];  // 1 or 2
  int r_pos[kMaxModes];  // position of R mode i in its source tensor
  int nk = 0;
  int k1[kMaxModes];  // contracted positions in T1 (ascending)
  int k2[kMaxModes];  // matching positions in T2
};

// Returns "" on success or an error message. `unsupported` is set when the
// labels are valid for NDTensors but outside this backend (too many modes,
// batch / trace labels).
std::string analyze_labels(int NR, const int64_t* lR, int N1, const int64_t* l1, int N2,
                           const int64_t* l2, Structure* st, bool* unsupported);

struct PairSpec {
  int64_t blk1, blk2, blkR;  // indices into the block lists
};

struct BlockList {
  int N = 0;
  std::vector<int32_t> dims;  // N per block
  std::vector<int64_t> offs;
  int64_t nnz = 0;
  int64_t nblk() const { return (int64_t)offs.size(); }
  const int32_t* d(int64_t b) const { return dims.data() + b * N; }
};

// Device-resident table (freed with the owner).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  hipError_t upload(const void* host, size_t n, hipStream_t s);
  // at least n bytes, contents undefined (frees and reallocates when short)
  hipError_t reserve(size_t n);
};

// Execution flags beyond the conjugation bits (values of include/itg.h's
// ITG_FLAG_KEEP_SLICES_1/2 and ITG_FLAG_REUSE_SLICES_1/2): split precision
// slices a fixed operand into storage owned by the plan and keeps it
// (KEEP), or uses what an earlier KEEP call left (REUSE) instead of slicing.
constexpr unsigned kFlagKeep1 = 4u, kFlagKeep2 = 8u, kFlagReuse1 = 16u, kFlagReuse2 = 32u;
constexpr unsigned kFlagConj = 3u;


struct PermuteLaunch {
  int nm = 0;
  int32_t perm[kMaxModes];
  int kind = 0;
  int64_t nblk = 0, tiles = 0;
  DevBuf blocks, prefix;
};

// Tile-shape choice "mixed": groups whose output block fills large tiles
// use the large shape, the others 64x64, as two grouped launches.
constexpr int kCfgMixed = 16;

struct GemmLaunch {
  int cfg = -1;
  int64_t grid = 0;  // workgroups, including idle padding slots
  DevBuf tiles;      // int2 {group, tile} per workgroup
  int cfg2 = -1;     // second launch of a mixed choice (64x64 part)
  int64_t grid2 = 0;
  DevBuf tiles2;
};

struct ExecPlan {
  enum Strategy { DIRECT, FIBER, GEMM } strategy = DIRECT;
  Structure st;
  // DIRECT
  int64_t d_ngroups = 0, d_tiles = 0;
  DevBuf d_pairs, d_groups, d_koffs, d_wgg;  // d_wgg: workgroup -> group
  // FIBER (A = the larger operand; f_a_is_2: A is T2)
  bool f_a_is_2 = false;
  bool f_mv = false;  // matrix form (fm_* tables) instead of the loop form
  int f_nf = 0;
  int f_smax = 0;     // matrix form: slot bound of the kernel (4, 8 or 16) >= every group's nslot
  int64_t f_wgs = 0;
  DevBuf f_groups, f_outs, f_pairs, f_koffs, f_gofs_c, f_gofs_b, f_wgg;
  DevBuf fm_groups, fm_slots, fm_outs, fm_coefs;
  // host copies of the matrix-form tables and the fiber modes (A mode, R
  // mode), for fusing two consecutive steps (build_fiber_chain)
  std::vector<FiberMVGroup> hfm_groups;
  std::vector<FiberSlot> hfm_slots;
  std::vector<FiberOutElem> hfm_outs;
  std::vector<FiberCoef> hfm_coefs;
  std::vector<std::pair<int, int>> f_fm;
  // GEMM
  bool swap = false, permA = false, permB = false, permC = false;
  int ta = 0, tb = 0;
  int64_t nnzA = 0, nnzB = 0, nnzC = 0;  // element counts of the (possibly permuted) operands
  PermuteLaunch pA, pB, pC;
  DevBuf g_pairs, g_groups;
  std::vector<GemmPair> h_pairs;
  std::vector<GemmGroup> h_groups;
  std::vector<double> group_cost;  // flops per group (for LPT ordering)
  std::map<int, std::unique_ptr<GemmLaunch>> launches;  // per tile cfg
  int tuned[2][3] = {{-1, -1, -1}, {-1, -1, -1}};      // autotuned cfg per (complex, prec)
  // split-K form (FP64): few output tiles over long contracted extents; the
  // K range of each output block is cut into slices whose partial sums go
  // to the workspace (ks_elems elements) and are added by the reduce kernel
  int64_t ks_grid = 0, ks_elems = 0, ks_total = 0, ks_maxk = 0;
  int ks_nred = 0, ks_maxs = 0;
  bool ks_dot = false;      // every output block has <= kKsDotMN elements
  bool ks_dot1 = false;     // ... and exactly one
  bool ks_skinny = false;   // every output block has a side <= kKsSkinny
  bool ks_kc = false;       // skinny, long side's rows contiguous along K
  int64_t ks_nrows = 0;     // skinny: workgroups (entry, first row)
  DevBuf ks_rows;
  int64_t ks_ngroups = 0;   // (block, slice) entries of the split tables
  DevBuf ks_pairs, ks_groups, ks_tiles, ks_red;
  double flops = 0.0;                                   // real flops (2*M*N*K summed)
  int64_t max_dim = 0;
  // split precision (real data): slice tables of the A-side operand P and
  // the B-side operand Q, read in place (no permutation), and the pair table
  bool split_ok = false;
  bool cplx_pair_alpha = false;  // some per-pair alpha has an imaginary part
  int64_t split_kmax = 0;        // largest contracted extent of a pair
  std::string split_why;         // reason when !split_ok
  int64_t s_blocks = 0, s_grid = 0;  // slicing grid over both operands
  int64_t sA_plane_bytes = 0, sB_plane_bytes = 0, sA_rows = 0, sB_rows = 0;
  int64_t sA_pmax = 0, sB_pmax = 0;  // partial row exponents (ints)
  DevBuf s_desc, s_pairs, s_groups;  // s_groups: h_groups over the K-chunked pairs
  // slicing grids: workgroup w of the launch over both operands slices tile
  // s_wgmap[w] = (block index into s_desc, tile index within the block); A's
  // blocks are s_desc[0, sA_blocks) and its tiles the first sA_grid entries;
  // a launch over B alone uses s_desc + sA_blocks with s_wgmapB (sB_grid
  // workgroups, block indices counted from B's first block)
  int64_t sA_blocks = 0, sA_grid = 0, sB_grid = 0;
  DevBuf s_wgmap, s_wgmapB;
  // digit planes and row exponents kept for a fixed operand (side 0 = A,
  // 1 = B) by ITG_FLAG_KEEP_SLICES_*: the source pointer they were made from
  // and the number of slice sets (1 real, 3 complex); 0 = none kept
  DevBuf kept[2];
  const void* kept_src[2] = {nullptr, nullptr};
  int kept_sets[2] = {0, 0};
};

// Build the execution plan of `pairs` (grouped: all pairs of one output block
// must be contiguous, in plan order). alpha per pair is taken from
// pair_alpha (2 doubles per pair) or 1.
// backend: kBackendAuto lets the shape decide (GEMM for a flop-weighted
// mean contracted extent >= 24, else FIBER or DIRECT); kBackendGemm takes
// the GEMM (TTGT on MFMA) whenever a mode is contracted; kBackendGeneric
// takes the generic strided kernel (DIRECT: no permutation, no MFMA) for
// every contraction.
constexpr int kBackendAuto = 0, kBackendGemm = 1, kBackendGeneric = 2;
std::string build_exec_plan(const Structure& st, const BlockList& b1, const BlockList& b2,
                            const BlockList& bR, const std::vector<PairSpec>& pairs,
                            const double* pair_alpha, hipStream_t s, ExecPlan* ep,
                            int backend = kBackendAuto);

// Slice and pair tables of the split-precision GEMM (sets split_ok / split_why).
std::string build_split_tables(ExecPlan* ep, const BlockList& P, const BlockList& Q,
                               const std::vector<int>& Mp, const std::vector<int>& Nq,
                               const std::vector<int>& Xp, const std::vector<int>& Xq,
                               const std::vector<PairSpec>& pairs, bool swap, hipStream_t s);

// Build the device tables of a grouped permutation (dst offsets default to
// the src offsets; a negative src offset marks an absent source block, whose
// values are zero). perm: new mode i = old mode perm[i], 0-based. fresh
// (optional, per block): 1 = the dst block is new, written without being read.
hipError_t build_permute_launch(PermuteLaunch* pl, int n, const int* perm, const BlockList& bl,
                                const std::vector<int64_t>* dst_offs, hipStream_t s,
                                const std::vector<uint8_t>* fresh = nullptr);

// Launch the plan. ws: workspace of at least exec_workspace_bytes().
// alpha_cplx: the call's alpha has a nonzero imaginary part.
size_t exec_workspace_bytes(const ExecPlan& ep, int cplx, int prec, bool alpha_cplx = false);
// kt (optional) brackets each launch with timer events. may_build = false:
// the launch tables of the tile shape must already exist (built by an
// earlier run with the same element type and precision); the call then
// allocates no device memory and fails with hipErrorNotReady otherwise.
hipError_t exec_run(ExecPlan& ep, int cplx, int prec, int cfg, hipStream_t s, void* ws,
                    const void* T1, const void* T2, void* R, unsigned flags, double a_re,
                    double a_im, double b_re, double b_im, std::string* err,
                    KernelTimer* kt = nullptr, bool may_build = true);
// Tile cfg picked by the built-in heuristic (the autotuner may override).
int default_gemm_cfg(const ExecPlan& ep, int cplx);
// Split-precision tile: 2 (128 x 64, three LDS stages, one workgroup per CU)
// when that shape still gives two workgroups per CU, else 0 (64 x 64).
int default_split_cfg(const ExecPlan& ep);
hipError_t ensure_gemm_launch(ExecPlan& ep, int cfg, int cplx, hipStream_t s,
                              bool may_build = true);
// XCD-balanced workgroup -> {group, tile} table for tile shape tc.
// filter: 0 = every group, 1 = only groups with M, N >= the tile, 2 = the rest.
std::vector<int2> xcd_tile_table(const ExecPlan& ep, GemmTileCfg tc, int filter = 0);

// Two consecutive matrix-form fiber contractions run as one pass (see
// ChainGroup). Built from the two steps' plans; the device tables live here.
struct ChainPlan {
  int nf = 0;            // modes of the fiber index (max of both steps)
  int xm = 0, ym = 0;    // unroll bounds (8 or 16) >= every group's nx / ny
  int64_t wgs = 0;
  DevBuf groups, slots, zouts, ymask, coefs1, coefs2, wgg;
};
// "" on success, "-" when the two steps cannot be fused (the caller runs
// them one after the other), otherwise an error message.
std::string build_fiber_chain(const ExecPlan& e1, const ExecPlan& e2, hipStream_t s,
                              ChainPlan* cp);
// Z = (X * B1) * B2 through a ChainPlan (unit scale, beta 0, no conjugation).
hipError_t exec_chain(const ChainPlan& cp, int cplx, hipStream_t s, const void* X, const void* B1,
                      const void* B2, void* Z, std::string* err, KernelTimer* kt = nullptr);

// ------------------------------------------------ execution without allocation
// itg_bs_execute on a caller-provided workspace of ws_bytes bytes: no
// device allocation (the plan's tables for these factors, element type and
// precision must exist already; ITG_ERR_INVALID otherwise, and when the
// contraction needs more workspace than ws_bytes). Float32 / ComplexF32 are
// refused (their widened copies would need memory). itg_api.cpp.
itg_status bs_execute_on(itg_handle* h, itg_bs_plan* p, itg_dtype dtype, void* R,
                         const void* T1, const void* T2, const double* alpha,
                         const double* pair_alpha, itg_prec prec, unsigned flags, void* ws,
                         size_t ws_bytes);
int handle_device(const itg_handle* h);
// Records msg as the handle's last error and returns st.
itg_status handle_fail(itg_handle* h, itg_status st, const char* msg);
// Per-device address of the library's fixed scratch buffer
// (itg_noalloc.hip): noalloc_setup looks it up once per device (with that
// device current; called by itg_init), noalloc_scratch returns it (nullptr
// before setup).
hipError_t noalloc_setup(int device);
void* noalloc_scratch(int device);

}  // namespace itg
