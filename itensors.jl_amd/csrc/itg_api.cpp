// C ABI of libitg.so (declared in include/itg.h).
//
// Dense and block-sparse contractions both lower to an ExecPlan (itg_plan.h):
// a dense call is a one-pair plan, cached per (dims, labels) so repeated calls
// with the same shapes (the steps of a network applied many times) skip
// planning and table uploads. A block-sparse plan is built once by
// itg_bs_plan_create and executed any number of times.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <unordered_map>

#include "../../include/itg.h"
#include "itg_plan.h"

using namespace itg;

static_assert(ITG_FLAG_KEEP_SLICES_1 == kFlagKeep1 && ITG_FLAG_KEEP_SLICES_2 == kFlagKeep2 &&
                  ITG_FLAG_REUSE_SLICES_1 == kFlagReuse1 && ITG_FLAG_REUSE_SLICES_2 == kFlagReuse2 &&
                  (ITG_FLAG_CONJ_A | ITG_FLAG_CONJ_B) == kFlagConj,
              "execution flag values");

struct itg_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* ws32 = nullptr;        // FP64 copies of Float32 / ComplexF32 operands
  size_t ws32_bytes = 0;
  std::vector<void*> retired;  // earlier workspaces, freed by itg_destroy
  bool timing = false;         // itg_timer_enable
  std::unique_ptr<KernelTimer> timer;
  std::vector<std::string> timer_names;  // snapshot behind itg_timer_read's `name`
  // dense plan cache (LRU)
  std::list<std::pair<std::string, std::unique_ptr<ExecPlan>>> dense_lru;
  std::unordered_map<std::string, decltype(dense_lru)::iterator> dense_map;
  // autotuned tile cfg per (bucket, cplx, prec)
  std::map<std::string, int> tuned;
  int split_tile = 0;  // itg_set_split_tile (0 = per plan)
  int backend = kBackendAuto;  // itg_set_gemm_backend: plans created from now on
  int bs_order = ITG_BS_ORDER_SEQUENTIAL;  // itg_set_blocksparse_order: likewise
  // itg_graph_begin / _end: the stream being captured, and the caller's
  // stream to restore (a handle on the null stream captures on its own)
  hipStream_t cap_stream = nullptr;  // created on first use, destroyed with the handle
  hipStream_t pre_capture = nullptr;
  bool capturing = false;
};

struct itg_graph {
  int device = 0;
  hipGraphExec_t exec = nullptr;
};

struct itg_bs_plan {
  int device = 0;
  int NR = 0;
  Structure st;
  BlockList b1, b2, bR;
  std::vector<int64_t> blocksR;  // NR x nblkR, 1-based
  std::vector<PairSpec> pairs;
  bool t2_outer = false;  // enumerated with T2's blocks outer (threaded order)
  std::unique_ptr<ExecPlan> ep;  // pair_alpha = NULL
  int backend = kBackendAuto;    // the handle's backend when the plan was created
  // one plan per set of per-pair factors used so far (the factors are baked
  // into the device tables). None is freed before the plan is destroyed: a
  // graph captured with one of them keeps reading its tables.
  std::map<std::vector<double>, std::unique_ptr<ExecPlan>> ep_alpha;
  // T2 block diagonal (itg_bs_diag_plan_create): device tables of
  // diag_contract_kernel instead of ep
  bool diag = false;
  int d_nr = 1;
  int64_t d_wgs = 0;
  DevBuf d_outs, d_pairs, d_wgg;
};

struct itg_bs_chain {
  int device = 0;
  ChainPlan cp;
};

struct itg_batch_plan {
  int device = 0;
  itg_dtype dtype = ITG_F64;
  const void* base1 = nullptr;
  const void* base2 = nullptr;
  void* baseR = nullptr;
  std::unique_ptr<ExecPlan> ep;
};

namespace {

itg_status fail(itg_handle* h, itg_status st, const std::string& msg) {
  if (h) h->err = msg;
  return st;
}

itg_status hip_fail(itg_handle* h, hipError_t e, const char* what) {
  return fail(h, e == hipErrorOutOfMemory ? ITG_ERR_NOMEM : ITG_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

bool dtype_ok(itg_dtype dt) { return dt == ITG_F64 || dt == ITG_C128; }
// Float32 / ComplexF32: accepted at the boundary; contractions run in FP64
// on widened copies and round once when R is written.
bool dtype32(itg_dtype dt) { return dt == ITG_F32 || dt == ITG_C64; }
bool dtype_any(itg_dtype dt) { return dtype_ok(dt) || dtype32(dt); }
itg_dtype wide_of(itg_dtype dt) { return dt == ITG_C64 ? ITG_C128 : ITG_F64; }
int elt_code(itg_dtype dt) {
  return dt == ITG_F64 ? kEltF64 : dt == ITG_C128 ? kEltC128 : dt == ITG_F32 ? kEltF32 : kEltC64;
}

// A grow-only device buffer: a captured graph keeps the address it was
// recorded with, so a buffer that has been handed to kernels is never freed
// before itg_destroy: it is retired and the new one at least doubles, which
// keeps the retired total below the live size.
hipError_t grow(itg_handle* h, void** buf, size_t* cur, size_t bytes) {
  if (bytes <= *cur) return hipSuccess;
  size_t b = std::max({bytes, (size_t)1 << 20, 2 * *cur});
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) return e;
  if (*buf) h->retired.push_back(*buf);
  *buf = p;
  *cur = b;
  return hipSuccess;
}

hipError_t ensure_ws(itg_handle* h, size_t bytes) { return grow(h, &h->ws, &h->ws_bytes, bytes); }

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// FP64 copies of up to two Float32 / ComplexF32 operands (n1, n2 elements)
// and an FP64 output of nR elements, carved from the handle's ws32.
struct Wide {
  void* t1 = nullptr;
  void* t2 = nullptr;
  void* r = nullptr;
};
itg_status widen_operands(itg_handle* h, itg_dtype dt, int64_t n1, const void* T1, int64_t n2,
                          const void* T2, int64_t nR, Wide* w) {
  const int cplx = dt == ITG_C64;
  const size_t es = cplx ? 16 : 8;
  const size_t b1 = al256((size_t)n1 * es), b2 = al256((size_t)n2 * es);
  hipError_t e = grow(h, &h->ws32, &h->ws32_bytes, b1 + b2 + al256((size_t)nR * es) + 256);
  if (e != hipSuccess) return hip_fail(h, e, "workspace allocation");
  char* p = (char*)h->ws32;
  w->t1 = p;
  w->t2 = p + b1;
  w->r = p + b1 + b2;
  if (T1 && (e = launch_widen(h->stream, cplx, n1, T1, w->t1)) != hipSuccess)
    return hip_fail(h, e, "widen launch");
  if (T2 && (e = launch_widen(h->stream, cplx, n2, T2, w->t2)) != hipSuccess)
    return hip_fail(h, e, "widen launch");
  return ITG_OK;
}

// One past the last element a block list's blocks occupy (diag: each block
// stores min(dims) elements).
int64_t extent_of(const BlockList& b, bool diag) {
  int64_t end = 0;
  for (int64_t k = 0; k < b.nblk(); ++k) {
    int64_t n = 1, mn = INT64_MAX;
    for (int i = 0; i < b.N; ++i) {
      n *= b.d(k)[i];
      mn = std::min<int64_t>(mn, b.d(k)[i]);
    }
    if (diag) n = b.N ? mn : 1;
    end = std::max(end, b.offs[k] + n);
  }
  return end;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

std::string tune_key(const ExecPlan& ep, int cplx, int prec) {
  // bucket by the log2 of the largest block dim and the transpose form
  int b = 0;
  while ((int64_t(1) << b) < ep.max_dim && b < 40) ++b;
  char buf[96];
  std::snprintf(buf, sizeof buf, "b%d_c%d_p%d_t%d%d", b, cplx, prec, ep.ta, ep.tb);
  return buf;
}

// The handle's timer while timing is on and the stream is not being captured.
KernelTimer* active_timer(itg_handle* h) {
  if (!h->timing || !h->timer) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(h->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return nullptr;
  return h->timer.get();
}

int pick_cfg(itg_handle* h, const ExecPlan& ep, int cplx, int prec) {
  if (prec != 0 && h->split_tile > 0) return h->split_tile == 1 ? 0 : h->split_tile;
  if (ep.tuned[cplx][prec] >= 0) return ep.tuned[cplx][prec];
  auto it = h->tuned.find(tune_key(ep, cplx, prec));
  if (it != h->tuned.end()) return it->second;
  return prec != 0 ? default_split_cfg(ep) : default_gemm_cfg(ep, cplx);
}

itg_status run_plan(itg_handle* h, ExecPlan& ep, int cplx, int prec, const void* T1,
                    const void* T2, void* R, unsigned flags, const double* alpha,
                    const double* beta) {
  const double a_re = alpha ? alpha[0] : 1.0, a_im = (alpha && cplx) ? alpha[1] : 0.0;
  const double b_re = beta ? beta[0] : 0.0, b_im = (beta && cplx) ? beta[1] : 0.0;
  hipError_t e = ensure_ws(h, exec_workspace_bytes(ep, cplx, prec, a_im != 0.0));
  if (e != hipSuccess) return hip_fail(h, e, "workspace allocation");
  std::string err;
  const int cfg = pick_cfg(h, ep, cplx, prec);
  e = exec_run(ep, cplx, prec, cfg, h->stream, h->ws, T1, T2, R, flags, a_re, a_im, b_re, b_im,
               &err, active_timer(h));
  if (e == hipErrorNotSupported) return fail(h, ITG_ERR_UNSUPPORTED, err);
  if (e != hipSuccess) return fail(h, ITG_ERR_HIP, err);
  return ITG_OK;
}

// ITG_FLAG_REUSE_SLICES_*: the plan must hold digit planes kept (by a KEEP
// call) from the same operand pointer with the same number of slice sets.
// Nothing to check outside split-precision GEMM plans, where the flags do
// nothing.
itg_status check_reuse(itg_handle* h, const ExecPlan& ep, int cplx, int prec, const void* T1,
                       const void* T2, unsigned flags) {
  if (prec == ITG_PREC_FP64 || ep.strategy != ExecPlan::GEMM || !ep.split_ok) return ITG_OK;
  for (int op = 0; op < 2; ++op) {
    const unsigned keep = op == 0 ? kFlagKeep1 : kFlagKeep2;
    const unsigned reuse = op == 0 ? kFlagReuse1 : kFlagReuse2;
    if (!(flags & reuse) || (flags & keep)) continue;
    const int side = ep.swap ? 1 - op : op;
    if (ep.kept_sets[side] != (cplx ? 3 : 1) || ep.kept_src[side] != (op == 0 ? T1 : T2))
      return fail(h, ITG_ERR_INVALID,
                  op == 0 ? "ITG_FLAG_REUSE_SLICES_1: this plan kept no slices of this T1"
                          : "ITG_FLAG_REUSE_SLICES_2: this plan kept no slices of this T2");
  }
  return ITG_OK;
}

// Enumeration order of a plan's pairs (they are stored grouped by output
// block): the order of the outer and inner loops over the block lists.
bool enum_before(const itg_bs_plan* p, const PairSpec& x, const PairSpec& y) {
  if (p->t2_outer) return x.blk2 != y.blk2 ? x.blk2 < y.blk2 : x.blk1 < y.blk1;
  return x.blk1 != y.blk1 ? x.blk1 < y.blk1 : x.blk2 < y.blk2;
}

// The execution plan of block-sparse plan p for the per-pair factors
// pair_alpha (NULL: the plan's own). The factors are baked into the device
// tables, so each new set builds (and keeps) a plan of its own; with build =
// false a set without tables is refused instead.
itg_status plan_for_factors(itg_handle* h, itg_bs_plan* p, const double* pair_alpha, int cplx,
                            bool build, ExecPlan** out) {
  *out = p->ep.get();
  if (!pair_alpha) return ITG_OK;
  // pair_alpha is given in enumeration order; the plan stores pairs grouped
  std::vector<size_t> idx(p->pairs.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    return enum_before(p, p->pairs[a], p->pairs[b]);
  });
  std::vector<double> pa(2 * p->pairs.size());
  for (size_t e = 0; e < idx.size(); ++e) {
    pa[2 * idx[e]] = pair_alpha[2 * e];
    pa[2 * idx[e] + 1] = cplx ? pair_alpha[2 * e + 1] : 0.0;
  }
  auto it = p->ep_alpha.find(pa);
  if (it == p->ep_alpha.end()) {
    if (!build)
      return fail(h, ITG_ERR_INVALID,
                  "these per-pair factors have no device tables yet: execute the plan once "
                  "with them through itg_bs_execute first");
    // building uploads tables on the stream, which a capture would record
    // with host sources that do not outlive the call
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
      return fail(h, ITG_ERR_INVALID,
                  "new per-pair factors during stream capture: execute once before capturing");
    std::unique_ptr<ExecPlan> np(new ExecPlan);
    std::string err =
        build_exec_plan(p->st, p->b1, p->b2, p->bR, p->pairs, pa.data(), h->stream, np.get(),
                        p->backend);
    if (!err.empty()) return fail(h, ITG_ERR_HIP, err);
    it = p->ep_alpha.emplace(pa, std::move(np)).first;
  }
  *out = it->second.get();
  return ITG_OK;
}

}  // namespace

namespace itg {

itg_status bs_execute_on(itg_handle* h, itg_bs_plan* p, itg_dtype dtype, void* R,
                         const void* T1, const void* T2, const double* alpha,
                         const double* pair_alpha, itg_prec prec, unsigned flags, void* ws,
                         size_t ws_bytes) {
  if (!h || !p) return ITG_ERR_INVALID;
  if (prec != ITG_PREC_FP64 && prec != ITG_PREC_SPLIT_BF16 && prec != ITG_PREC_SPLIT_INT8)
    return fail(h, ITG_ERR_INVALID, "unknown precision mode");
  if (dtype32(dtype))
    return fail(h, ITG_ERR_UNSUPPORTED,
                "Float32 / ComplexF32 operands need widened copies: use itg_bs_execute");
  if (!dtype_ok(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (p->diag) return fail(h, ITG_ERR_INVALID, "diagonal plan (use itg_bs_diag_execute)");
  if (p->pairs.empty()) return ITG_OK;
  if (!R || !T1 || !T2) return fail(h, ITG_ERR_INVALID, "null data pointer");
  if (p->device != h->device) return fail(h, ITG_ERR_INVALID, "plan built on another device");
  if (!ws) return fail(h, ITG_ERR_INVALID, "no workspace");
  const int cplx = dtype == ITG_C128;
  DeviceGuard g(h->device);
  if (flags & (kFlagKeep1 | kFlagKeep2))
    return fail(h, ITG_ERR_INVALID,
                "ITG_FLAG_KEEP_SLICES_* allocates the plan's slice storage: use itg_bs_execute");
  ExecPlan* ep = nullptr;
  itg_status st = plan_for_factors(h, p, pair_alpha, cplx, false, &ep);
  if (st != ITG_OK) return st;
  if ((st = check_reuse(h, *ep, cplx, (int)prec, T1, T2, flags)) != ITG_OK) return st;
  const double a_re = alpha ? alpha[0] : 1.0, a_im = (alpha && cplx) ? alpha[1] : 0.0;
  const size_t need = exec_workspace_bytes(*ep, cplx, (int)prec, a_im != 0.0);
  if (need > ws_bytes) {
    char msg[200];
    std::snprintf(msg, sizeof msg,
                  "this contraction needs %zu bytes of workspace, above the %zu bytes of the "
                  "library's fixed scratch buffer: use itg_bs_execute",
                  need, ws_bytes);
    return fail(h, ITG_ERR_INVALID, msg);
  }
  std::string err;
  const int cfg = pick_cfg(h, *ep, cplx, (int)prec);
  hipError_t e = exec_run(*ep, cplx, (int)prec, cfg, h->stream, ws, T1, T2, R, flags, a_re, a_im,
                          0.0, 0.0, &err, nullptr, false);
  if (e == hipErrorNotSupported) return fail(h, ITG_ERR_UNSUPPORTED, err);
  if (e == hipErrorNotReady) return fail(h, ITG_ERR_INVALID, err);
  if (e != hipSuccess) return fail(h, ITG_ERR_HIP, err);
  return ITG_OK;
}

int handle_device(const itg_handle* h) { return h->device; }

itg_status handle_fail(itg_handle* h, itg_status st, const char* msg) {
  return fail(h, st, msg);
}

}  // namespace itg

extern "C" {

const char* itg_version(void) { return "0.1.0"; }

itg_status itg_init(int device, itg_handle** out) {
  if (!out) return ITG_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return ITG_ERR_HIP;
  if (device < 0 || device >= n) return ITG_ERR_INVALID;
  // the address of this device's copy of the fixed scratch buffer
  // (itg_bs_execute_noalloc), resolved once per device
  if (noalloc_setup(device) != hipSuccess) return ITG_ERR_HIP;
  itg_handle* h = new itg_handle;
  h->device = device;
  *out = h;
  return ITG_OK;
}

itg_status itg_destroy(itg_handle* h) {
  if (!h) return ITG_OK;
  {
    DeviceGuard g(h->device);
    if (h->capturing) {
      hipGraph_t gr = nullptr;
      (void)hipStreamEndCapture(h->stream, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      h->stream = h->pre_capture;
    }
    if (h->ws || !h->retired.empty()) (void)hipStreamSynchronize(h->stream);
    if (h->ws) (void)hipFree(h->ws);
    if (h->ws32) (void)hipFree(h->ws32);
    for (void* p : h->retired) (void)hipFree(p);
    h->dense_map.clear();
    h->dense_lru.clear();
    if (h->cap_stream) {
      (void)hipStreamSynchronize(h->cap_stream);
      (void)hipStreamDestroy(h->cap_stream);
    }
  }
  delete h;
  return ITG_OK;
}

itg_status itg_graph_begin(itg_handle* h) {
  if (!h) return ITG_ERR_INVALID;
  if (h->capturing) return fail(h, ITG_ERR_INVALID, "a capture is already open on this handle");
  DeviceGuard g(h->device);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (h->stream && (hipStreamIsCapturing(h->stream, &cs) != hipSuccess ||
                    cs != hipStreamCaptureStatusNone))
    return fail(h, ITG_ERR_INVALID, "the handle's stream is already being captured");
  h->pre_capture = h->stream;
  if (!h->stream) {
    // the null stream cannot be captured: record on the handle's own stream,
    // ordered after the work already queued on the null stream
    if (!h->cap_stream) {
      hipError_t e = hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking);
      if (e != hipSuccess) return hip_fail(h, e, "capture stream");
    }
    hipError_t e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(h, e, "hipStreamSynchronize");
    h->stream = h->cap_stream;
  }
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    h->stream = h->pre_capture;
    return hip_fail(h, e, "hipStreamBeginCapture");
  }
  h->capturing = true;
  return ITG_OK;
}

itg_status itg_graph_end(itg_handle* h, itg_graph** out) {
  if (!h || !out) return ITG_ERR_INVALID;
  *out = nullptr;
  if (!h->capturing) return fail(h, ITG_ERR_INVALID, "no capture is open on this handle");
  DeviceGuard g(h->device);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(h->stream, &gr);
  h->stream = h->pre_capture;
  h->capturing = false;
  if (e != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    return hip_fail(h, e, "capture failed (a call inside it needed a new plan table, workspace "
                          "or per-pair factors: run it once before capturing)");
  }
  std::unique_ptr<itg_graph> ig(new itg_graph);
  ig->device = h->device;
  e = hipGraphInstantiate(&ig->exec, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (e != hipSuccess) return hip_fail(h, e, "hipGraphInstantiate");
  *out = ig.release();
  return ITG_OK;
}

itg_status itg_graph_launch(itg_handle* h, itg_graph* gph) {
  if (!h || !gph) return ITG_ERR_INVALID;
  if (gph->device != h->device) return fail(h, ITG_ERR_INVALID, "graph built on another device");
  if (h->capturing) return fail(h, ITG_ERR_INVALID, "cannot launch a graph while capturing");
  DeviceGuard g(h->device);
  hipError_t e = hipGraphLaunch(gph->exec, h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "hipGraphLaunch");
  return ITG_OK;
}

itg_status itg_graph_destroy(itg_graph* gph) {
  if (!gph) return ITG_OK;
  DeviceGuard g(gph->device);
  (void)hipDeviceSynchronize();
  if (gph->exec) (void)hipGraphExecDestroy(gph->exec);
  delete gph;
  return ITG_OK;
}

const char* itg_last_error(const itg_handle* h) { return h ? h->err.c_str() : "null handle"; }

itg_status itg_set_stream(itg_handle* h, void* stream) {
  if (!h) return ITG_ERR_INVALID;
  if (h->capturing) {
    // naming the stream the capture started from keeps the capture's stream
    if ((hipStream_t)stream == h->pre_capture) return ITG_OK;
    return fail(h, ITG_ERR_INVALID, "cannot change the stream while a capture is open");
  }
  h->stream = (hipStream_t)stream;
  return ITG_OK;
}

itg_status itg_synchronize(itg_handle* h) {
  if (!h) return ITG_ERR_INVALID;
  DeviceGuard g(h->device);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "hipStreamSynchronize");
  return ITG_OK;
}

itg_status itg_set_split_tile(itg_handle* h, int tile) {
  if (!h || tile < 0 || tile > 3) return ITG_ERR_INVALID;
  h->split_tile = tile;
  return ITG_OK;
}

itg_status itg_set_gemm_backend(itg_handle* h, int backend) {
  if (!h || backend < kBackendAuto || backend > kBackendGeneric) return ITG_ERR_INVALID;
  h->backend = backend;
  return ITG_OK;
}

itg_status itg_set_blocksparse_order(itg_handle* h, int order) {
  if (!h || (order != ITG_BS_ORDER_SEQUENTIAL && order != ITG_BS_ORDER_THREADED))
    return ITG_ERR_INVALID;
  h->bs_order = order;
  return ITG_OK;
}

itg_status itg_timer_enable(itg_handle* h, int enable) {
  if (!h) return ITG_ERR_INVALID;
  DeviceGuard g(h->device);
  if (enable) {
    if (!h->timer) h->timer.reset(new KernelTimer);
    h->timer->clear();
  }
  h->timing = enable != 0;
  return ITG_OK;
}

int64_t itg_timer_count(itg_handle* h) {
  if (!h || !h->timer) return 0;
  DeviceGuard g(h->device);
  (void)h->timer->fold();
  return (int64_t)h->timer->totals.size();
}

itg_status itg_timer_read(itg_handle* h, int64_t i, const char** name, int64_t* calls,
                          double* total_ms) {
  if (!h || !name || !calls || !total_ms) return ITG_ERR_INVALID;
  if (!h->timer) return fail(h, ITG_ERR_INVALID, "timer entry out of range");
  DeviceGuard g(h->device);
  hipError_t e = h->timer->fold();
  if (e != hipSuccess) return hip_fail(h, e, "timer events");
  if (i < 0 || i >= (int64_t)h->timer->totals.size())
    return fail(h, ITG_ERR_INVALID, "timer entry out of range");
  h->timer_names.clear();
  for (const auto& kv : h->timer->totals) h->timer_names.push_back(kv.first);
  auto it = h->timer->totals.find(h->timer_names[(size_t)i]);
  *name = h->timer_names[(size_t)i].c_str();
  *calls = it->second.first;
  *total_ms = it->second.second;
  return ITG_OK;
}

itg_status itg_malloc(itg_handle* h, size_t bytes, void** ptr) {
  if (!h || !ptr) return ITG_ERR_INVALID;
  *ptr = nullptr;
  if (bytes == 0) return ITG_OK;
  DeviceGuard g(h->device);
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return hip_fail(h, e, "hipMalloc");
  return ITG_OK;
}

itg_status itg_free(itg_handle* h, void* ptr) {
  if (!h) return ITG_ERR_INVALID;
  if (!ptr) return ITG_OK;
  DeviceGuard g(h->device);
  // queued kernels may still use the buffer
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipFree(ptr);
  if (e != hipSuccess) return hip_fail(h, e, "hipFree");
  return ITG_OK;
}

static itg_status copy_bytes(itg_handle* h, void* dst, const void* src, size_t bytes,
                             hipMemcpyKind kind, bool sync) {
  if (!h) return ITG_ERR_INVALID;
  if (bytes == 0) return ITG_OK;
  if (!dst || !src) return fail(h, ITG_ERR_INVALID, "null pointer");
  DeviceGuard g(h->device);
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, h->stream);
  if (e == hipSuccess && sync) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "hipMemcpyAsync");
  return ITG_OK;
}

itg_status itg_memcpy_htod(itg_handle* h, void* dst, const void* src, size_t bytes) {
  return copy_bytes(h, dst, src, bytes, hipMemcpyHostToDevice, true);
}
itg_status itg_memcpy_dtoh(itg_handle* h, void* dst, const void* src, size_t bytes) {
  return copy_bytes(h, dst, src, bytes, hipMemcpyDeviceToHost, true);
}
itg_status itg_convert(itg_handle* h, itg_dtype from, const void* src, itg_dtype to, void* dst,
                       int64_t n) {
  if (!h) return ITG_ERR_INVALID;
  if (!dtype_any(from) || !dtype_any(to)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (n < 0) return fail(h, ITG_ERR_INVALID, "negative length");
  if (n == 0) return ITG_OK;
  if (!src || !dst) return fail(h, ITG_ERR_INVALID, "null data pointer");
  DeviceGuard g(h->device);
  if (from == to) {
    const size_t es = from == ITG_C128 ? 16 : (from == ITG_F32 ? 4 : 8);
    hipError_t e = hipMemcpyAsync(dst, src, (size_t)n * es, hipMemcpyDeviceToDevice, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "hipMemcpyAsync");
    return ITG_OK;
  }
  hipError_t e = launch_convert(h->stream, elt_code(from), elt_code(to), n, src, dst);
  if (e == hipErrorInvalidValue)
    return fail(h, ITG_ERR_INVALID, "not a widening conversion (Float32 < Float64, real < complex)");
  if (e != hipSuccess) return hip_fail(h, e, "convert launch");
  return ITG_OK;
}

itg_status itg_memcpy_dtod(itg_handle* h, void* dst, const void* src, size_t bytes) {
  return copy_bytes(h, dst, src, bytes, hipMemcpyDeviceToDevice, false);
}

itg_status itg_permute(itg_handle* h, itg_dtype dtype, int ndim, const int64_t* dims,
                       const int32_t* perm, const void* src, void* dst, const double* alpha,
                       const double* beta) {
  if (!h) return ITG_ERR_INVALID;
  if (!dtype_any(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (ndim < 0 || ndim > kMaxModes) return fail(h, ITG_ERR_UNSUPPORTED, "too many modes");
  if (ndim > 0 && (!dims || !perm)) return fail(h, ITG_ERR_INVALID, "null dims/perm");
  int p0[kMaxModes];
  bool seen[kMaxModes] = {false};
  int64_t n = 1;
  for (int i = 0; i < ndim; ++i) {
    if (dims[i] < 0) return fail(h, ITG_ERR_INVALID, "negative dimension");
    n *= dims[i];
    const int q = perm[i] - 1;
    if (q < 0 || q >= ndim || seen[q]) return fail(h, ITG_ERR_INVALID, "perm is not a permutation");
    seen[q] = true;
    p0[i] = q;
  }
  if (n > INT32_MAX) return fail(h, ITG_ERR_UNSUPPORTED, "tensor too large");
  if (n == 0) return ITG_OK;
  if (!src || !dst) return fail(h, ITG_ERR_INVALID, "null data pointer");
  DeviceGuard g(h->device);
  // one-block grouped permutation
  BlockList bl;
  bl.N = ndim;
  for (int i = 0; i < ndim; ++i) bl.dims.push_back((int32_t)dims[i]);
  bl.offs.push_back(0);
  bl.nnz = n;
  PermuteLaunch pl;
  {
    hipError_t e = build_permute_launch(&pl, ndim, p0, bl, nullptr, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "permute tables");
  }
  const int cplx = dtype == ITG_C128 || dtype == ITG_C64;
  const double a_re = alpha ? alpha[0] : 1.0, a_im = (alpha && cplx) ? alpha[1] : 0.0;
  const double b_re = beta ? beta[0] : 0.0, b_im = (beta && cplx) ? beta[1] : 0.0;
  hipError_t e = launch_permute(h->stream, elt_code(dtype), pl.nm, pl.perm, (const PermBlock*)pl.blocks.p,
                                (const int64_t*)pl.prefix.p, pl.nblk, pl.tiles, pl.kind, src, dst,
                                a_re, a_im, b_re, b_im);
  if (e != hipSuccess) return hip_fail(h, e, "permute launch");
  // the tables are freed when `pl` goes out of scope: wait for the kernel
  e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "permute");
  return ITG_OK;
}

itg_status itg_contract_dense_ex(itg_handle* h, itg_dtype dtype, int NR, const int64_t* dimsR,
                                 const int64_t* labelsR, void* R, int N1, const int64_t* dims1,
                                 const int64_t* labels1, const void* T1, int N2,
                                 const int64_t* dims2, const int64_t* labels2, const void* T2,
                                 const double* alpha, const double* beta, itg_prec prec,
                                 unsigned flags) {
  if (!h) return ITG_ERR_INVALID;
  flags &= kFlagConj;  // the slice-keeping flags are for block-sparse plans only
  if (!dtype_any(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (prec != ITG_PREC_FP64 && prec != ITG_PREC_SPLIT_BF16 && prec != ITG_PREC_SPLIT_INT8)
    return fail(h, ITG_ERR_INVALID, "unknown precision mode");
  if ((NR > 0 && (!dimsR || !labelsR)) || (N1 > 0 && (!dims1 || !labels1)) ||
      (N2 > 0 && (!dims2 || !labels2)))
    return fail(h, ITG_ERR_INVALID, "null dims/labels");
  Structure st;
  bool unsup = false;
  std::string m = analyze_labels(NR, labelsR, N1, labels1, N2, labels2, &st, &unsup);
  if (!m.empty()) return fail(h, unsup ? ITG_ERR_UNSUPPORTED : ITG_ERR_INVALID, m);
  // dims consistency
  int64_t n1 = 1, n2 = 1, nR = 1;
  for (int i = 0; i < N1; ++i) {
    if (dims1[i] < 0) return fail(h, ITG_ERR_INVALID, "negative dimension");
    n1 *= dims1[i];
  }
  for (int i = 0; i < N2; ++i) {
    if (dims2[i] < 0) return fail(h, ITG_ERR_INVALID, "negative dimension");
    n2 *= dims2[i];
  }
  for (int r = 0; r < NR; ++r) {
    const int64_t want = st.r_src[r] == 1 ? dims1[st.r_pos[r]] : dims2[st.r_pos[r]];
    if (dimsR[r] != want) return fail(h, ITG_ERR_INVALID, "dimension mismatch between R and inputs");
    nR *= dimsR[r];
  }
  for (int i = 0; i < st.nk; ++i)
    if (dims1[st.k1[i]] != dims2[st.k2[i]])
      return fail(h, ITG_ERR_INVALID, "contracted dimensions do not match");
  if (n1 > INT32_MAX || n2 > INT32_MAX || nR > INT32_MAX)
    return fail(h, ITG_ERR_UNSUPPORTED, "tensor with more than 2^31-1 elements");
  if (nR == 0) return ITG_OK;
  if (!R || (n1 > 0 && !T1) || (n2 > 0 && !T2)) return fail(h, ITG_ERR_INVALID, "null data pointer");
  if (dtype32(dtype)) {
    // FP64 (MFMA) on widened copies, whatever the precision mode: Float32
    // values are exact in FP64, so the split form has nothing to emulate
    DeviceGuard g(h->device);
    Wide w;
    itg_status st = widen_operands(h, dtype, n1, T1, n2, T2, nR, &w);
    if (st != ITG_OK) return st;
    st = itg_contract_dense_ex(h, wide_of(dtype), NR, dimsR, labelsR, w.r, N1, dims1, labels1,
                               w.t1, N2, dims2, labels2, w.t2, alpha, nullptr, ITG_PREC_FP64,
                               flags);
    if (st != ITG_OK) return st;
    const int c = dtype == ITG_C64;
    hipError_t e = launch_narrow(h->stream, c, nR, w.r, R, beta ? beta[0] : 0.0,
                                 (beta && c) ? beta[1] : 0.0);
    if (e != hipSuccess) return hip_fail(h, e, "narrow launch");
    return ITG_OK;
  }
  const int cplx = dtype == ITG_C128;
  DeviceGuard g(h->device);

  // plan cache key
  std::string key;
  key.reserve(256);
  auto app = [&](const int64_t* v, int n) {
    key.append((const char*)&n, sizeof n);
    key.append((const char*)v, sizeof(int64_t) * (size_t)n);
  };
  app(dimsR, NR);
  app(labelsR, NR);
  app(dims1, N1);
  app(labels1, N1);
  app(dims2, N2);
  app(labels2, N2);
  key.append((const char*)&h->backend, sizeof h->backend);
  ExecPlan* ep = nullptr;
  auto it = h->dense_map.find(key);
  if (it != h->dense_map.end()) {
    h->dense_lru.splice(h->dense_lru.begin(), h->dense_lru, it->second);
    ep = it->second->second.get();
  } else {
    BlockList b1, b2, bR;
    b1.N = N1;
    b2.N = N2;
    bR.N = NR;
    for (int i = 0; i < N1; ++i) b1.dims.push_back((int32_t)dims1[i]);
    for (int i = 0; i < N2; ++i) b2.dims.push_back((int32_t)dims2[i]);
    for (int i = 0; i < NR; ++i) bR.dims.push_back((int32_t)dimsR[i]);
    if (N1 == 0) b1.dims.push_back(1);
    if (N2 == 0) b2.dims.push_back(1);
    if (NR == 0) bR.dims.push_back(1);
    b1.offs = {0};
    b2.offs = {0};
    bR.offs = {0};
    b1.nnz = n1;
    b2.nnz = n2;
    bR.nnz = nR;
    std::vector<PairSpec> pairs{{0, 0, 0}};
    std::unique_ptr<ExecPlan> np(new ExecPlan);
    if (n1 == 0 || n2 == 0) {
      // nothing to sum: R = beta*R, done by a K = 0 direct pass
    }
    std::string err =
        build_exec_plan(st, b1, b2, bR, pairs, nullptr, h->stream, np.get(), h->backend);
    if (!err.empty()) return fail(h, ITG_ERR_HIP, err);
    h->dense_lru.emplace_front(key, std::move(np));
    h->dense_map[key] = h->dense_lru.begin();
    ep = h->dense_lru.front().second.get();
    while (h->dense_lru.size() > 64) {
      // evicted plans' tables may still be read by queued kernels
      (void)hipStreamSynchronize(h->stream);
      h->dense_map.erase(h->dense_lru.back().first);
      h->dense_lru.pop_back();
    }
  }
  return run_plan(h, *ep, cplx, (int)prec, T1, T2, R, flags, alpha, beta);
}

itg_status itg_contract_dense(itg_handle* h, itg_dtype dtype, int NR, const int64_t* dimsR,
                              const int64_t* labelsR, void* R, int N1, const int64_t* dims1,
                              const int64_t* labels1, const void* T1, int N2,
                              const int64_t* dims2, const int64_t* labels2, const void* T2,
                              const double* alpha, const double* beta, itg_prec prec) {
  return itg_contract_dense_ex(h, dtype, NR, dimsR, labelsR, R, N1, dims1, labels1, T1, N2, dims2,
                               labels2, T2, alpha, beta, prec, 0u);
}

// ----------------------------------------------------------- block sparse
}  // extern "C"

namespace {
// Block lists, output structure and the pair list of T1 x T2
// (contract_blockoffsets, NDTensors/src/blocksparse/contract_sequential.jl:1-41);
// shared by the plain and the diagonal plans. `diag2`: T2 is block diagonal
// (every block (k, ..., k)) and each block stores only its diagonal, min of
// its dims elements. `t2_outer`: enumerate T2's blocks in the outer loop and
// T1's in the inner one (the threaded order when T1 has no more blocks than
// T2, contract_threaded.jl:45-72); otherwise T1's outer (contract_sequential.jl:26-39,
// and contract_threaded.jl:14-44 for T1 with more blocks).
itg_status build_bs_structure(itg_handle* h, int N1, int64_t nblk1, const int64_t* blocks1,
                              const int64_t* offs1, const int64_t* nsec1,
                              const int64_t* secdims1, const int64_t* labels1, int N2,
                              int64_t nblk2, const int64_t* blocks2, const int64_t* offs2,
                              const int64_t* nsec2, const int64_t* secdims2,
                              const int64_t* labels2, int NR, const int64_t* labelsR,
                              bool diag2, bool t2_outer, std::unique_ptr<itg_bs_plan>& p) {
  if (nblk1 < 0 || nblk2 < 0) return fail(h, ITG_ERR_INVALID, "negative block count");
  if ((nblk1 > 0 && (!blocks1 || !offs1)) || (nblk2 > 0 && (!blocks2 || !offs2)))
    return fail(h, ITG_ERR_INVALID, "null block arrays");
  if ((N1 > 0 && (!nsec1 || !secdims1 || !labels1)) || (N2 > 0 && (!nsec2 || !secdims2 || !labels2)) ||
      (NR > 0 && !labelsR))
    return fail(h, ITG_ERR_INVALID, "null sector/label arrays");
  Structure st;
  bool unsup = false;
  std::string m = analyze_labels(NR, labelsR, N1, labels1, N2, labels2, &st, &unsup);
  if (!m.empty()) return fail(h, unsup ? ITG_ERR_UNSUPPORTED : ITG_ERR_INVALID, m);
  // sector tables
  std::vector<std::vector<int64_t>> sec1(N1), sec2(N2);
  {
    int64_t o = 0;
    for (int i = 0; i < N1; ++i) {
      if (nsec1[i] < 0) return fail(h, ITG_ERR_INVALID, "negative sector count");
      sec1[i].assign(secdims1 + o, secdims1 + o + nsec1[i]);
      o += nsec1[i];
    }
    o = 0;
    for (int i = 0; i < N2; ++i) {
      if (nsec2[i] < 0) return fail(h, ITG_ERR_INVALID, "negative sector count");
      sec2[i].assign(secdims2 + o, secdims2 + o + nsec2[i]);
      o += nsec2[i];
    }
  }
  for (int i = 0; i < st.nk; ++i)
    if (sec1[st.k1[i]] != sec2[st.k2[i]])
      return fail(h, ITG_ERR_INVALID, "contracted indices have different block structure");
  p.reset(new itg_bs_plan);
  p->device = h->device;
  p->NR = NR;
  p->st = st;
  auto fill = [&](BlockList& bl, int N, int64_t nblk, const int64_t* blocks, const int64_t* offs,
                  const std::vector<std::vector<int64_t>>& sec, bool diag) -> std::string {
    bl.N = N;
    bl.dims.resize((size_t)std::max<int64_t>(nblk, 0) * std::max(N, 1), 1);
    bl.offs.assign(offs, offs + nblk);
    int64_t end = 0;
    for (int64_t b = 0; b < nblk; ++b) {
      int64_t sz = 1, dl = INT64_MAX;
      for (int i = 0; i < N; ++i) {
        const int64_t c = blocks[b * N + i];
        if (c < 1 || c > (int64_t)sec[i].size()) return "block coordinate out of range";
        bl.dims[b * std::max(N, 1) + i] = (int32_t)sec[i][c - 1];
        sz *= sec[i][c - 1];
        dl = std::min(dl, sec[i][c - 1]);
        if (diag && c != blocks[b * N]) return "the diagonal tensor must be block diagonal";
      }
      if (sz > INT32_MAX) return "block with more than 2^31-1 elements";
      if (offs[b] < 0) return "negative block offset";
      end = std::max(end, offs[b] + (diag ? (N > 0 ? dl : 1) : sz));
    }
    bl.nnz = end;
    return "";
  };
  std::string err = fill(p->b1, N1, nblk1, blocks1, offs1, sec1, false);
  if (err.empty()) err = fill(p->b2, N2, nblk2, blocks2, offs2, sec2, diag2);
  if (!err.empty()) return fail(h, ITG_ERR_INVALID, err);

  // pair enumeration: the inner operand's blocks are bucketed by their
  // contracted coordinates (in block order), so the scan is
  // O(nblk1 + nblk2 + npairs) while keeping the (outer, inner) order of the
  // nested loops over the two block lists
  std::unordered_map<std::string, std::vector<int64_t>> by_key;
  auto kkey = [&](const int64_t* blk, const int* pos) {
    std::string k;
    for (int i = 0; i < st.nk; ++i) k.append((const char*)&blk[pos[i]], sizeof(int64_t));
    return k;
  };
  if (t2_outer) {
    for (int64_t b = 0; b < nblk1; ++b) by_key[kkey(blocks1 + b * N1, st.k1)].push_back(b);
  } else {
    for (int64_t b = 0; b < nblk2; ++b) by_key[kkey(blocks2 + b * N2, st.k2)].push_back(b);
  }
  std::map<std::vector<int64_t>, int64_t> rindex;  // block coords -> R block index
  BlockList& bR = p->bR;
  bR.N = NR;
  int64_t nnzR = 0;
  const int64_t nouter = t2_outer ? nblk2 : nblk1;
  for (int64_t bo = 0; bo < nouter; ++bo) {
    auto f = t2_outer ? by_key.find(kkey(blocks2 + bo * N2, st.k2))
                      : by_key.find(kkey(blocks1 + bo * N1, st.k1));
    if (f == by_key.end()) continue;
    for (int64_t bi : f->second) {
      const int64_t b1 = t2_outer ? bi : bo, b2 = t2_outer ? bo : bi;
      std::vector<int64_t> cR(NR);
      for (int r = 0; r < NR; ++r)
        cR[r] = st.r_src[r] == 1 ? blocks1[b1 * N1 + st.r_pos[r]] : blocks2[b2 * N2 + st.r_pos[r]];
      auto ri = rindex.find(cR);
      int64_t idx;
      if (ri == rindex.end()) {
        idx = (int64_t)bR.offs.size();
        rindex.emplace(cR, idx);
        int64_t sz = 1;
        for (int r = 0; r < NR; ++r) {
          const int64_t d = st.r_src[r] == 1 ? sec1[st.r_pos[r]][cR[r] - 1] : sec2[st.r_pos[r]][cR[r] - 1];
          bR.dims.push_back((int32_t)d);
          sz *= d;
        }
        if (NR == 0) bR.dims.push_back(1);
        bR.offs.push_back(nnzR);
        nnzR += sz;
        for (int r = 0; r < NR; ++r) p->blocksR.push_back(cR[r]);
      } else {
        idx = ri->second;
      }
      p->pairs.push_back(PairSpec{b1, b2, idx});
    }
  }
  bR.nnz = nnzR;
  // group pairs by output block (stable: plan order within a block), as
  // contract_generic.jl:57-60 does
  std::stable_sort(p->pairs.begin(), p->pairs.end(),
                   [](const PairSpec& a, const PairSpec& b) { return a.blkR < b.blkR; });
  if (N1 == 0) p->b1.dims.assign((size_t)nblk1, 1);
  if (N2 == 0) p->b2.dims.assign((size_t)nblk2, 1);
  return ITG_OK;
}

}  // namespace

extern "C" {

itg_status itg_bs_plan_create(itg_handle* h, int N1, int64_t nblk1, const int64_t* blocks1,
                              const int64_t* offs1, const int64_t* nsec1,
                              const int64_t* secdims1, const int64_t* labels1, int N2,
                              int64_t nblk2, const int64_t* blocks2, const int64_t* offs2,
                              const int64_t* nsec2, const int64_t* secdims2,
                              const int64_t* labels2, int NR, const int64_t* labelsR,
                              itg_bs_plan** out) {
  if (!h || !out) return ITG_ERR_INVALID;
  *out = nullptr;
  std::unique_ptr<itg_bs_plan> p;
  // the threaded algorithm takes T2's blocks outer unless T1 has more
  // blocks (contract_threaded.jl:14, 45)
  const bool t2_outer = h->bs_order == ITG_BS_ORDER_THREADED && !(nblk1 > nblk2);
  itg_status st = build_bs_structure(h, N1, nblk1, blocks1, offs1, nsec1, secdims1, labels1, N2,
                                     nblk2, blocks2, offs2, nsec2, secdims2, labels2, NR,
                                     labelsR, false, t2_outer, p);
  if (st != ITG_OK) return st;
  p->t2_outer = t2_outer;
  if (!p->pairs.empty()) {
    DeviceGuard g(h->device);
    p->backend = h->backend;
    p->ep.reset(new ExecPlan);
    std::string err = build_exec_plan(p->st, p->b1, p->b2, p->bR, p->pairs, nullptr, h->stream,
                                      p->ep.get(), p->backend);
    if (!err.empty()) return fail(h, ITG_ERR_HIP, err);
  }
  *out = p.release();
  return ITG_OK;
}

itg_status itg_bs_diag_plan_create(itg_handle* h, int N1, int64_t nblk1,
                                   const int64_t* blocks1, const int64_t* offs1,
                                   const int64_t* nsec1, const int64_t* secdims1,
                                   const int64_t* labels1, int ND, int64_t nblkD,
                                   const int64_t* blocksD, const int64_t* offsD,
                                   const int64_t* nsecD, const int64_t* secdimsD,
                                   const int64_t* labelsD, int NR, const int64_t* labelsR,
                                   itg_bs_plan** out) {
  if (!h || !out) return ITG_ERR_INVALID;
  *out = nullptr;
  std::unique_ptr<itg_bs_plan> p;
  itg_status st = build_bs_structure(h, N1, nblk1, blocks1, offs1, nsec1, secdims1, labels1, ND,
                                     nblkD, blocksD, offsD, nsecD, secdimsD, labelsD, NR,
                                     labelsR, true, false, p);
  if (st != ITG_OK) return st;
  p->diag = true;
  p->d_nr = std::max(NR, 1);
  std::vector<DiagOut> outs;
  std::vector<DiagPair> dps;
  std::vector<int32_t> wgg;
  int64_t wgs = 0;
  const Structure& S = p->st;
  for (size_t q = 0; q < p->pairs.size(); ++q) {
    const PairSpec& ps = p->pairs[q];
    if (outs.empty() || q == 0 || p->pairs[q - 1].blkR != ps.blkR) {
      DiagOut o;
      std::memset(&o, 0, sizeof(o));
      o.offR = p->bR.offs[(size_t)ps.blkR];
      o.nelem = 1;
      for (int i = 0; i < kMaxModes; ++i) o.dims[i] = 1;
      for (int r = 0; r < NR; ++r) {
        o.dims[r] = p->bR.d(ps.blkR)[r];
        o.nelem *= o.dims[r];
      }
      o.p0 = o.p1 = (int32_t)q;
      o.wg0 = wgs;
      const int64_t n = (o.nelem + kDiagChunk - 1) / kDiagChunk;
      for (int64_t w = 0; w < n; ++w) wgg.push_back((int32_t)outs.size());
      wgs += n;
      outs.push_back(o);
    }
    outs.back().p1 = (int32_t)q + 1;
    int64_t s1[kMaxModes];
    {
      int64_t a = 1;  // column-major strides of the T1 block
      for (int i = 0; i < p->b1.N; ++i) {
        s1[i] = a;
        a *= p->b1.d(ps.blk1)[i];
      }
    }
    DiagPair d;
    std::memset(&d, 0, sizeof(d));
    d.off1 = p->b1.offs[(size_t)ps.blk1];
    d.offD = p->b2.offs[(size_t)ps.blk2];
    d.dlen = ND > 0 ? INT32_MAX : 1;
    for (int i = 0; i < ND; ++i) d.dlen = std::min(d.dlen, p->b2.d(ps.blk2)[i]);
    for (int i = 0; i < S.nk; ++i) d.kstr += s1[S.k1[i]];
    d.u0 = -1;
    for (int r = 0; r < NR; ++r) {
      if (S.r_src[r] == 1) {
        d.str1[r] = (int32_t)s1[S.r_pos[r]];
      } else if (d.u0 < 0) {
        d.u0 = r;
      } else {
        d.umask |= 1u << r;
      }
    }
    dps.push_back(d);
  }
  if (!outs.empty()) {
    DeviceGuard g(h->device);
    hipError_t e;
    if ((e = p->d_outs.upload(outs.data(), outs.size() * sizeof(DiagOut), h->stream)) !=
            hipSuccess ||
        (e = p->d_pairs.upload(dps.data(), dps.size() * sizeof(DiagPair), h->stream)) !=
            hipSuccess ||
        (e = p->d_wgg.upload(wgg.data(), wgg.size() * sizeof(int32_t), h->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
      return hip_fail(h, e, "diagonal plan table upload");
  }
  p->d_wgs = wgs;
  *out = p.release();
  return ITG_OK;
}

itg_status itg_bs_diag_execute(itg_handle* h, itg_bs_plan* p, itg_dtype dtype, void* R,
                               const void* T1, const void* D, const double* uniform,
                               const double* alpha) {
  if (!h || !p) return ITG_ERR_INVALID;
  if (!p->diag) return fail(h, ITG_ERR_INVALID, "not a diagonal plan (use itg_bs_execute)");
  if (!dtype_any(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (p->pairs.empty()) return ITG_OK;
  if (!R || !T1 || (!D && !uniform)) return fail(h, ITG_ERR_INVALID, "null data pointer");
  if (p->device != h->device) return fail(h, ITG_ERR_INVALID, "plan built on another device");
  if (dtype32(dtype)) {
    DeviceGuard g(h->device);
    Wide w;
    itg_status st = widen_operands(h, dtype, extent_of(p->b1, false), T1,
                                   uniform ? 0 : extent_of(p->b2, true), uniform ? nullptr : D,
                                   p->bR.nnz, &w);
    if (st != ITG_OK) return st;
    st = itg_bs_diag_execute(h, p, wide_of(dtype), w.r, w.t1, uniform ? nullptr : w.t2, uniform,
                             alpha);
    if (st != ITG_OK) return st;
    hipError_t e = launch_narrow(h->stream, dtype == ITG_C64, p->bR.nnz, w.r, R, 0.0, 0.0);
    if (e != hipSuccess) return hip_fail(h, e, "narrow launch");
    return ITG_OK;
  }
  const int cplx = dtype == ITG_C128;
  DeviceGuard g(h->device);
  const double a_re = alpha ? alpha[0] : 1.0, a_im = (alpha && cplx) ? alpha[1] : 0.0;
  const double u_re = uniform ? uniform[0] : 0.0, u_im = (uniform && cplx) ? uniform[1] : 0.0;
  KernelTimer* kt = active_timer(h);
  if (kt) kt->begin("diag_contract_kernel", h->stream);
  hipError_t e = launch_diag(h->stream, cplx, p->d_nr, (const DiagOut*)p->d_outs.p,
                             (const DiagPair*)p->d_pairs.p, (const int32_t*)p->d_wgg.p, p->d_wgs,
                             T1, uniform ? T1 : D, uniform ? 1 : 0, u_re, u_im, a_re, a_im, R);
  if (kt) kt->end(h->stream);
  if (e != hipSuccess)
    return fail(h, ITG_ERR_HIP, std::string("diagonal kernel launch: ") + hipGetErrorString(e));
  return ITG_OK;
}

itg_status itg_bs_plan_sizes(const itg_bs_plan* p, int64_t* nblkR, int64_t* nnzR,
                             int64_t* npairs) {
  if (!p) return ITG_ERR_INVALID;
  if (nblkR) *nblkR = p->bR.nblk();
  if (nnzR) *nnzR = p->bR.nnz;
  if (npairs) *npairs = (int64_t)p->pairs.size();
  return ITG_OK;
}

itg_status itg_bs_plan_info(const itg_bs_plan* p, int32_t* info) {
  if (!p || !info) return ITG_ERR_INVALID;
  info[0] = info[1] = info[2] = info[3] = 0;
  if (!p->ep || p->diag) {
    info[0] = -1;
    return ITG_OK;
  }
  const ExecPlan& ep = *p->ep;
  info[0] = ep.strategy == ExecPlan::GEMM ? 2 : ep.strategy == ExecPlan::FIBER ? 1 : 0;
  if (ep.strategy == ExecPlan::GEMM) {
    // permA / permB name the GEMM's operands, which are T2 / T1 when swapped
    info[1] = (ep.swap ? ep.permB : ep.permA) ? 1 : 0;
    info[2] = (ep.swap ? ep.permA : ep.permB) ? 1 : 0;
    info[3] = ep.permC ? 1 : 0;
  }
  return ITG_OK;
}

itg_status itg_bs_plan_blocks(const itg_bs_plan* p, int64_t* blocksR, int64_t* offsR) {
  if (!p) return ITG_ERR_INVALID;
  if (blocksR && !p->blocksR.empty())
    std::memcpy(blocksR, p->blocksR.data(), p->blocksR.size() * sizeof(int64_t));
  if (offsR && !p->bR.offs.empty())
    std::memcpy(offsR, p->bR.offs.data(), p->bR.offs.size() * sizeof(int64_t));
  return ITG_OK;
}

itg_status itg_bs_plan_pairs(const itg_bs_plan* p, int64_t* pairs) {
  if (!p) return ITG_ERR_INVALID;
  // report in enumeration order
  std::vector<PairSpec> v(p->pairs);
  std::stable_sort(v.begin(), v.end(), [p](const PairSpec& a, const PairSpec& b) {
    return enum_before(p, a, b);
  });
  for (size_t i = 0; i < v.size(); ++i) {
    pairs[3 * i] = v[i].blk1;
    pairs[3 * i + 1] = v[i].blk2;
    pairs[3 * i + 2] = v[i].blkR;
  }
  return ITG_OK;
}

itg_status itg_bs_execute(itg_handle* h, itg_bs_plan* p, itg_dtype dtype, void* R,
                          const void* T1, const void* T2, const double* alpha,
                          const double* pair_alpha, itg_prec prec, unsigned flags) {
  if (!h || !p) return ITG_ERR_INVALID;
  if (!dtype_any(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (prec != ITG_PREC_FP64 && prec != ITG_PREC_SPLIT_BF16 && prec != ITG_PREC_SPLIT_INT8)
    return fail(h, ITG_ERR_INVALID, "unknown precision mode");
  if (p->diag) return fail(h, ITG_ERR_INVALID, "diagonal plan (use itg_bs_diag_execute)");
  if (p->pairs.empty()) return ITG_OK;
  if (!R || !T1 || !T2) return fail(h, ITG_ERR_INVALID, "null data pointer");
  if (p->device != h->device) return fail(h, ITG_ERR_INVALID, "plan built on another device");
  if (dtype32(dtype)) {
    DeviceGuard g(h->device);
    Wide w;
    itg_status st = widen_operands(h, dtype, extent_of(p->b1, false), T1,
                                   extent_of(p->b2, false), T2, p->bR.nnz, &w);
    if (st != ITG_OK) return st;
    st = itg_bs_execute(h, p, wide_of(dtype), w.r, w.t1, w.t2, alpha, pair_alpha,
                        ITG_PREC_FP64, flags & kFlagConj);
    if (st != ITG_OK) return st;
    hipError_t e = launch_narrow(h->stream, dtype == ITG_C64, p->bR.nnz, w.r, R, 0.0, 0.0);
    if (e != hipSuccess) return hip_fail(h, e, "narrow launch");
    return ITG_OK;
  }
  const int cplx = dtype == ITG_C128;
  DeviceGuard g(h->device);
  ExecPlan* ep = nullptr;
  itg_status st = plan_for_factors(h, p, pair_alpha, cplx, true, &ep);
  if (st != ITG_OK) return st;
  if ((st = check_reuse(h, *ep, cplx, (int)prec, T1, T2, flags)) != ITG_OK) return st;
  return run_plan(h, *ep, cplx, (int)prec, T1, T2, R, flags, alpha, nullptr);
}

itg_status itg_bs_plan_destroy(itg_bs_plan* p) {
  if (!p) return ITG_OK;
  DeviceGuard g(p->device);
  (void)hipDeviceSynchronize();
  delete p;
  return ITG_OK;
}

itg_status itg_bs_chain_create(itg_handle* h, const itg_bs_plan* p1, const itg_bs_plan* p2,
                               itg_bs_chain** out) {
  if (!h || !p1 || !p2 || !out) return ITG_ERR_INVALID;
  *out = nullptr;
  if (p1->device != h->device || p2->device != h->device)
    return fail(h, ITG_ERR_INVALID, "plan built on another device");
  const BlockList &y1 = p1->bR, &y2 = p2->b1;
  if (y1.N != y2.N || y1.offs != y2.offs || y1.dims != y2.dims)
    return fail(h, ITG_ERR_INVALID,
                "the first plan's output structure is not the second plan's first operand");
  if (p1->pairs.empty() || p2->pairs.empty() || !p1->ep || !p2->ep)
    return fail(h, ITG_ERR_UNSUPPORTED, "empty or diagonal plan: no fused form");
  DeviceGuard g(h->device);
  std::unique_ptr<itg_bs_chain> c(new itg_bs_chain);
  c->device = h->device;
  const std::string err = build_fiber_chain(*p1->ep, *p2->ep, h->stream, &c->cp);
  if (err == "-")
    return fail(h, ITG_ERR_UNSUPPORTED, "the two contractions have no fused form");
  if (!err.empty()) return fail(h, ITG_ERR_HIP, err);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "chain table upload");
  *out = c.release();
  return ITG_OK;
}

itg_status itg_bs_chain_execute(itg_handle* h, itg_bs_chain* c, itg_dtype dtype, void* Z,
                                const void* X, const void* B1, const void* B2) {
  if (!h || !c) return ITG_ERR_INVALID;
  if (!dtype_ok(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (!Z || !X || !B1 || !B2) return fail(h, ITG_ERR_INVALID, "null data pointer");
  if (c->device != h->device) return fail(h, ITG_ERR_INVALID, "chain built on another device");
  DeviceGuard g(h->device);
  std::string err;
  hipError_t e = exec_chain(c->cp, dtype == ITG_C128, h->stream, X, B1, B2, Z, &err,
                            active_timer(h));
  if (e != hipSuccess) return fail(h, ITG_ERR_HIP, err);
  return ITG_OK;
}

itg_status itg_bs_chain_destroy(itg_bs_chain* c) {
  if (!c) return ITG_OK;
  DeviceGuard g(c->device);
  (void)hipDeviceSynchronize();
  delete c;
  return ITG_OK;
}

itg_status itg_bs_permute_ex(itg_handle* h, itg_dtype dtype, int N, int64_t nblk,
                             const int64_t* blockdims, const int64_t* offs_src,
                             const int64_t* offs_dst, const uint8_t* fresh, const int32_t* perm,
                             const void* src, void* dst, const double* alpha, const double* beta,
                             int32_t* nonzero) {
  if (!h) return ITG_ERR_INVALID;
  if (!dtype_any(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (N < 0 || N > kMaxModes) return fail(h, ITG_ERR_UNSUPPORTED, "too many modes");
  if (nblk < 0) return fail(h, ITG_ERR_INVALID, "negative block count");
  if (nblk == 0) return ITG_OK;
  if (!blockdims || !offs_src || !offs_dst || (N > 0 && !perm) || !dst)
    return fail(h, ITG_ERR_INVALID, "null argument");
  int p0[kMaxModes];
  bool seen[kMaxModes] = {false};
  for (int i = 0; i < N; ++i) {
    const int q = perm[i] - 1;
    if (q < 0 || q >= N || seen[q]) return fail(h, ITG_ERR_INVALID, "perm is not a permutation");
    seen[q] = true;
    p0[i] = q;
  }
  BlockList bl;
  bl.N = N;
  bool any_src = false;
  for (int64_t b = 0; b < nblk; ++b) {
    int64_t sz = 1;
    for (int i = 0; i < N; ++i) {
      if (blockdims[b * N + i] < 0) return fail(h, ITG_ERR_INVALID, "negative dimension");
      bl.dims.push_back((int32_t)blockdims[b * N + i]);
      sz *= blockdims[b * N + i];
    }
    if (sz > INT32_MAX) return fail(h, ITG_ERR_UNSUPPORTED, "block too large");
    if (N == 0) bl.dims.push_back(1);
    if (offs_dst[b] < 0) return fail(h, ITG_ERR_INVALID, "negative destination offset");
    any_src |= offs_src[b] >= 0;
  }
  if (any_src && !src) return fail(h, ITG_ERR_INVALID, "null source");
  bl.offs.assign(offs_src, offs_src + nblk);
  std::vector<int64_t> od(offs_dst, offs_dst + nblk);
  std::vector<uint8_t> fr;
  if (fresh) fr.assign(fresh, fresh + nblk);
  DeviceGuard g(h->device);
  PermuteLaunch pl;
  hipError_t e = build_permute_launch(&pl, N, p0, bl, &od, h->stream, fresh ? &fr : nullptr);
  if (e != hipSuccess) return hip_fail(h, e, "permute tables");
  DevBuf nz;
  if (nonzero) {
    std::vector<int32_t> zero((size_t)nblk, 0);
    if ((e = nz.upload(zero.data(), zero.size() * sizeof(int32_t), h->stream)) != hipSuccess)
      return hip_fail(h, e, "flag buffer");
  }
  const int cplx = dtype == ITG_C128 || dtype == ITG_C64;
  const double a_re = alpha ? alpha[0] : 1.0, a_im = (alpha && cplx) ? alpha[1] : 0.0;
  const double b_re = beta ? beta[0] : 0.0, b_im = (beta && cplx) ? beta[1] : 0.0;
  KernelTimer* kt = active_timer(h);
  if (kt) kt->begin("permute_kernel", h->stream);
  e = launch_permute(h->stream, elt_code(dtype), pl.nm, pl.perm, (const PermBlock*)pl.blocks.p,
                     (const int64_t*)pl.prefix.p, pl.nblk, pl.tiles, pl.kind, src, dst, a_re,
                     a_im, b_re, b_im, nonzero ? (int32_t*)nz.p : nullptr);
  if (kt) kt->end(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "permute launch");
  if (nonzero) {
    e = hipMemcpyAsync(nonzero, nz.p, (size_t)nblk * sizeof(int32_t), hipMemcpyDeviceToHost,
                       h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "flag download");
  }
  // the tables (and the flags) are freed when the locals go out of scope
  e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "permute");
  return ITG_OK;
}

itg_status itg_bs_permute(itg_handle* h, itg_dtype dtype, int N, int64_t nblk,
                          const int64_t* blockdims, const int64_t* offs_src,
                          const int64_t* offs_dst, const int32_t* perm, const void* src,
                          void* dst) {
  if (!h) return ITG_ERR_INVALID;
  if (nblk > 0 && offs_src)
    for (int64_t b = 0; b < nblk; ++b)
      if (offs_src[b] < 0) return fail(h, ITG_ERR_INVALID, "negative source offset");
  return itg_bs_permute_ex(h, dtype, N, nblk, blockdims, offs_src, offs_dst, nullptr, perm, src,
                           dst, nullptr, nullptr, nullptr);
}

itg_status itg_bs_copy(itg_handle* h, itg_dtype dtype, int64_t nblk, int nmodes,
                       const int64_t* dims, const int64_t* src_off, const int64_t* src_str,
                       const int64_t* dst_off, const int64_t* dst_str, const void* src,
                       int64_t src_len, void* dst, int64_t dst_len) {
  if (!h) return ITG_ERR_INVALID;
  if (!dtype_ok(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (nmodes < 0 || nmodes > kMaxModes) return fail(h, ITG_ERR_UNSUPPORTED, "too many modes");
  if (nblk < 0) return fail(h, ITG_ERR_INVALID, "negative block count");
  if (nblk == 0) return ITG_OK;
  if (!src_off || !dst_off || !src || !dst || (nmodes > 0 && (!dims || !src_str || !dst_str)))
    return fail(h, ITG_ERR_INVALID, "null argument");
  std::vector<CopyBlock> cb((size_t)nblk);
  std::vector<int64_t> prefix(1, 0);
  for (int64_t b = 0; b < nblk; ++b) {
    CopyBlock& d = cb[(size_t)b];
    std::memset(&d, 0, sizeof(d));
    const int64_t* db = dims + b * nmodes;
    const int64_t* sb = src_str + b * nmodes;
    const int64_t* tb = dst_str + b * nmodes;
    // traversal order: the destination's smallest stride first (coalesced stores)
    int ord[kMaxModes];
    for (int i = 0; i < nmodes; ++i) ord[i] = i;
    std::stable_sort(ord, ord + nmodes, [&](int x, int y) {
      return std::llabs(tb[x]) < std::llabs(tb[y]);
    });
    int64_t n = 1, slo = src_off[b], shi = src_off[b], dlo = dst_off[b], dhi = dst_off[b];
    for (int i = 0; i < nmodes; ++i) {
      const int m = ord[i];
      if (db[m] < 0) return fail(h, ITG_ERR_INVALID, "negative dimension");
      d.dims[i] = (int32_t)std::min<int64_t>(db[m], INT32_MAX);
      d.sstr[i] = sb[m];
      d.dstr[i] = tb[m];
      n *= db[m];
      if (n > INT32_MAX) return fail(h, ITG_ERR_UNSUPPORTED, "block too large");
      if (db[m] > 0) {
        const int64_t es = (db[m] - 1) * sb[m], ed = (db[m] - 1) * tb[m];
        (es < 0 ? slo : shi) += es;
        (ed < 0 ? dlo : dhi) += ed;
      }
    }
    for (int i = nmodes; i < kMaxModes; ++i) d.dims[i] = 1;
    d.off_src = src_off[b];
    d.off_dst = dst_off[b];
    d.nelem = n;
    if (n > 0 && (slo < 0 || shi >= src_len || dlo < 0 || dhi >= dst_len))
      return fail(h, ITG_ERR_INVALID, "block copy reaches outside the source or destination");
    prefix.push_back(prefix.back() + copy_block_wgs(n));
  }
  DeviceGuard g(h->device);
  DevBuf tb, tp;
  hipError_t e;
  if ((e = tb.upload(cb.data(), cb.size() * sizeof(CopyBlock), h->stream)) != hipSuccess ||
      (e = tp.upload(prefix.data(), prefix.size() * sizeof(int64_t), h->stream)) != hipSuccess)
    return hip_fail(h, e, "copy tables");
  KernelTimer* kt = active_timer(h);
  if (kt) kt->begin("copy_blocks_kernel", h->stream);
  e = launch_copy_blocks(h->stream, dtype == ITG_C128, std::max(nmodes, 1), (const CopyBlock*)tb.p,
                         (const int64_t*)tp.p, nblk, prefix.back(), src, dst);
  if (kt) kt->end(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "copy launch");
  e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "copy");
  return ITG_OK;
}

itg_status itg_bs_autotune(itg_handle* h, itg_bs_plan* p, itg_dtype dtype, void* R,
                           const void* T1, const void* T2, itg_prec prec) {
  if (!h || !p) return ITG_ERR_INVALID;
  if (!dtype_ok(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (p->pairs.empty() || !p->ep || p->ep->strategy != ExecPlan::GEMM) return ITG_OK;
  const int cplx = dtype == ITG_C128;
  DeviceGuard g(h->device);
  ExecPlan& ep = *p->ep;
  if (prec != ITG_PREC_FP64 && !ep.split_ok) return ITG_OK;
  hipError_t e = ensure_ws(h, exec_workspace_bytes(ep, cplx, (int)prec));
  if (e != hipSuccess) return hip_fail(h, e, "workspace allocation");
  hipEvent_t t0, t1;
  if ((e = hipEventCreate(&t0)) != hipSuccess) return hip_fail(h, e, "event");
  if ((e = hipEventCreate(&t1)) != hipSuccess) {
    (void)hipEventDestroy(t0);
    return hip_fail(h, e, "event");
  }
  int best = -1;
  float best_ms = 1e30f;
  std::string err;
  std::vector<int> cands;
  if (prec != ITG_PREC_FP64) {
    cands = {0, 2};  // split tiles 64 x 64 and 128 x 64
    if (!cplx) cands.push_back(3);  // 64 x 64 with A in registers
  } else {
    for (int c = 0; c < gemm_num_cfgs(); ++c)
      if (gemm_cfg_ok(c, cplx)) cands.push_back(c);
    cands.push_back(kCfgMixed);
  }
  for (int cfg : cands) {
    // warm-up, then 3 timed runs
    e = exec_run(ep, cplx, (int)prec, cfg, h->stream, h->ws, T1, T2, R, 0u, 1.0, 0.0, 0.0, 0.0, &err);
    if (e != hipSuccess) break;
    (void)hipEventRecord(t0, h->stream);
    for (int r = 0; r < 3; ++r)
      (void)exec_run(ep, cplx, (int)prec, cfg, h->stream, h->ws, T1, T2, R, 0u, 1.0, 0.0, 0.0, 0.0, &err);
    (void)hipEventRecord(t1, h->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t0, t1);
    if (ms < best_ms) {
      best_ms = ms;
      best = cfg;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (e == hipErrorNotSupported) return fail(h, ITG_ERR_UNSUPPORTED, err);
  if (e != hipSuccess) return fail(h, ITG_ERR_HIP, err);
  if (best >= 0) {
    h->tuned[tune_key(ep, cplx, (int)prec)] = best;  // for plans not tuned themselves
    ep.tuned[cplx][(int)prec] = best;
  }
  return ITG_OK;
}

// ------------------------------------------------------------- batches
itg_status itg_batch_plan_create(itg_handle* h, itg_dtype dtype, int64_t nbatch, int NR,
                                 const int64_t* dimsR, const int64_t* labelsR, void* const* R,
                                 int N1, const int64_t* dims1, const int64_t* labels1,
                                 const void* const* T1, int N2, const int64_t* dims2,
                                 const int64_t* labels2, const void* const* T2,
                                 itg_batch_plan** out) {
  if (!h || !out) return ITG_ERR_INVALID;
  *out = nullptr;
  if (!dtype_ok(dtype)) return fail(h, ITG_ERR_UNSUPPORTED, "unsupported dtype");
  if (nbatch <= 0) return fail(h, ITG_ERR_INVALID, "empty batch");
  if (!R || !T1 || !T2 || (NR > 0 && (!dimsR || !labelsR)) || (N1 > 0 && (!dims1 || !labels1)) ||
      (N2 > 0 && (!dims2 || !labels2)))
    return fail(h, ITG_ERR_INVALID, "null argument");
  Structure st;
  bool unsup = false;
  std::string m = analyze_labels(NR, labelsR, N1, labels1, N2, labels2, &st, &unsup);
  if (!m.empty()) return fail(h, unsup ? ITG_ERR_UNSUPPORTED : ITG_ERR_INVALID, m);
  const int64_t es = dtype == ITG_C128 ? 16 : 8;
  std::unique_ptr<itg_batch_plan> p(new itg_batch_plan);
  p->device = h->device;
  p->dtype = dtype;
  // base pointer of each operand = lowest item address; items become
  // element offsets from it
  auto base_of = [&](const void* const* ptrs) {
    uintptr_t b = UINTPTR_MAX;
    for (int64_t i = 0; i < nbatch; ++i) b = std::min(b, (uintptr_t)ptrs[i]);
    return b;
  };
  const uintptr_t b1 = base_of(T1), b2 = base_of(T2), bR = base_of((const void* const*)R);
  p->base1 = (const void*)b1;
  p->base2 = (const void*)b2;
  p->baseR = (void*)bR;
  BlockList L1, L2, LR;
  L1.N = N1;
  L2.N = N2;
  LR.N = NR;
  std::vector<PairSpec> pairs;
  for (int64_t i = 0; i < nbatch; ++i) {
    const int64_t* d1 = dims1 + i * N1;
    const int64_t* d2 = dims2 + i * N2;
    const int64_t* dR = dimsR + i * NR;
    int64_t n1 = 1, n2 = 1, nR = 1;
    for (int k = 0; k < N1; ++k) {
      if (d1[k] < 0) return fail(h, ITG_ERR_INVALID, "negative dimension");
      n1 *= d1[k];
      L1.dims.push_back((int32_t)d1[k]);
    }
    for (int k = 0; k < N2; ++k) {
      if (d2[k] < 0) return fail(h, ITG_ERR_INVALID, "negative dimension");
      n2 *= d2[k];
      L2.dims.push_back((int32_t)d2[k]);
    }
    for (int r = 0; r < NR; ++r) {
      const int64_t want = st.r_src[r] == 1 ? d1[st.r_pos[r]] : d2[st.r_pos[r]];
      if (dR[r] != want) return fail(h, ITG_ERR_INVALID, "dimension mismatch between R and inputs");
      nR *= dR[r];
      LR.dims.push_back((int32_t)dR[r]);
    }
    for (int k = 0; k < st.nk; ++k)
      if (d1[st.k1[k]] != d2[st.k2[k]])
        return fail(h, ITG_ERR_INVALID, "contracted dimensions do not match");
    if (n1 > INT32_MAX || n2 > INT32_MAX || nR > INT32_MAX)
      return fail(h, ITG_ERR_UNSUPPORTED, "item with more than 2^31-1 elements");
    const uintptr_t a1 = (uintptr_t)T1[i], a2 = (uintptr_t)T2[i], aR = (uintptr_t)R[i];
    if ((a1 - b1) % es || (a2 - b2) % es || (aR - bR) % es)
      return fail(h, ITG_ERR_INVALID, "item pointers are not element-aligned");
    if (N1 == 0) L1.dims.push_back(1);
    if (N2 == 0) L2.dims.push_back(1);
    if (NR == 0) LR.dims.push_back(1);
    L1.offs.push_back((int64_t)((a1 - b1) / es));
    L2.offs.push_back((int64_t)((a2 - b2) / es));
    LR.offs.push_back((int64_t)((aR - bR) / es));
    L1.nnz = std::max(L1.nnz, L1.offs.back() + n1);
    L2.nnz = std::max(L2.nnz, L2.offs.back() + n2);
    LR.nnz = std::max(LR.nnz, LR.offs.back() + nR);
    pairs.push_back(PairSpec{i, i, i});
  }
  DeviceGuard g(h->device);
  p->ep.reset(new ExecPlan);
  m = build_exec_plan(st, L1, L2, LR, pairs, nullptr, h->stream, p->ep.get(), h->backend);
  if (!m.empty()) return fail(h, ITG_ERR_HIP, m);
  *out = p.release();
  return ITG_OK;
}

itg_status itg_batch_execute(itg_handle* h, itg_batch_plan* p, const double* alpha,
                             itg_prec prec) {
  if (!h || !p) return ITG_ERR_INVALID;
  if (prec != ITG_PREC_FP64 && prec != ITG_PREC_SPLIT_BF16 && prec != ITG_PREC_SPLIT_INT8)
    return fail(h, ITG_ERR_INVALID, "unknown precision mode");
  if (p->device != h->device) return fail(h, ITG_ERR_INVALID, "plan built on another device");
  DeviceGuard g(h->device);
  return run_plan(h, *p->ep, p->dtype == ITG_C128, (int)prec, p->base1, p->base2, p->baseR, 0u,
                  alpha, nullptr);
}

itg_status itg_batch_plan_destroy(itg_batch_plan* p) {
  if (!p) return ITG_OK;
  DeviceGuard g(p->device);
  (void)hipDeviceSynchronize();
  delete p;
  return ITG_OK;
}

}  // extern "C"
