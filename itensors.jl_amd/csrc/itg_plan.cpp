// Host-side contraction planner of libitg.so. See itg_plan.h.
#include "itg_plan.h"

#include <algorithm>
#include <map>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <numeric>

namespace itg {

namespace {
// Error of a run that may not build tables (exec_run with may_build = false).
constexpr const char* kTablesMissing =
    "the launch tables of this tile shape are not built yet: execute the plan once with "
    "itg_bs_execute in the same element type and precision first";
}  // namespace

DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

hipError_t DevBuf::reserve(size_t n) {
  if (n <= bytes) return hipSuccess;
  if (p) {
    hipError_t e = hipFree(p);
    if (e != hipSuccess) return e;
    p = nullptr;
    bytes = 0;
  }
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) return e;
  bytes = n;
  return hipSuccess;
}

hipError_t DevBuf::upload(const void* host, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (n > bytes) {
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) return e;
    bytes = n;
  }
  // the host staging vectors die when the planner returns: wait for the copy
  hipError_t e = hipMemcpyAsync(p, host, n, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

// ------------------------------------------------------------- labels
// Label rules of NDTensors.contract (contraction_logic.jl:5-62): a label
// present in both inputs is contracted; every other label must appear once in
// the output, which holds exactly the uncontracted labels.
std::string analyze_labels(int NR, const int64_t* lR, int N1, const int64_t* l1, int N2,
                           const int64_t* l2, Structure* st, bool* unsupported) {
  *unsupported = false;
  if (NR < 0 || N1 < 0 || N2 < 0) return "negative number of modes";
  if (NR > kMaxModes || N1 > kMaxModes || N2 > kMaxModes) {
    *unsupported = true;
    return "more than 8 modes per tensor is not supported by this backend";
  }
  for (int i = 0; i < N1; ++i)
    for (int j = i + 1; j < N1; ++j)
      if (l1[i] == l1[j]) {
        *unsupported = true;
        return "repeated label within T1 (trace) is not supported";
      }
  for (int i = 0; i < N2; ++i)
    for (int j = i + 1; j < N2; ++j)
      if (l2[i] == l2[j]) {
        *unsupported = true;
        return "repeated label within T2 (trace) is not supported";
      }
  for (int i = 0; i < NR; ++i)
    for (int j = i + 1; j < NR; ++j)
      if (lR[i] == lR[j]) return "repeated label in R";
  Structure s;
  s.NR = NR;
  s.N1 = N1;
  s.N2 = N2;
  s.nk = 0;
  int pos1R[kMaxModes], pos2R[kMaxModes];
  for (int i = 0; i < N1; ++i) {
    int in2 = -1, inR = -1;
    for (int j = 0; j < N2; ++j)
      if (l2[j] == l1[i]) in2 = j;
    for (int j = 0; j < NR; ++j)
      if (lR[j] == l1[i]) inR = j;
    if (in2 >= 0 && inR >= 0) {
      *unsupported = true;
      return "a label shared by T1, T2 and R (batch mode) is not supported";
    }
    if (in2 >= 0) {
      s.k1[s.nk] = i;
      s.k2[s.nk] = in2;
      s.nk++;
      pos1R[i] = -1;
    } else {
      if (inR < 0) return "a free label of T1 is missing from R";
      pos1R[i] = inR;
    }
  }
  for (int j = 0; j < N2; ++j) {
    bool in1 = false;
    for (int i = 0; i < N1; ++i)
      if (l1[i] == l2[j]) in1 = true;
    int inR = -1;
    for (int r = 0; r < NR; ++r)
      if (lR[r] == l2[j]) inR = r;
    if (in1) {
      pos2R[j] = -1;
    } else {
      if (inR < 0) return "a free label of T2 is missing from R";
      pos2R[j] = inR;
    }
  }
  if (NR != N1 + N2 - 2 * s.nk) return "R labels do not match the uncontracted labels";
  for (int r = 0; r < NR; ++r) s.r_src[r] = 0;
  for (int i = 0; i < N1; ++i)
    if (pos1R[i] >= 0) {
      s.r_src[pos1R[i]] = 1;
      s.r_pos[pos1R[i]] = i;
    }
  for (int j = 0; j < N2; ++j)
    if (pos2R[j] >= 0) {
      s.r_src[pos2R[j]] = 2;
      s.r_pos[pos2R[j]] = j;
    }
  for (int r = 0; r < NR; ++r)
    if (s.r_src[r] == 0) return "an R label is not a free label of T1 or T2";
  *st = s;
  return "";
}

// --------------------------------------------------------- permutations
namespace {

// Merge src modes that stay adjacent (and in order) under perm. Returns the
// merged permutation and, for each merged mode, the range of original modes.
struct Merged {
  int nm = 0;
  int32_t perm[kMaxModes];
  int first[kMaxModes], last[kMaxModes];  // original src modes of merged src mode
};

Merged merge_modes(int n, const int* perm) {
  // group dst positions into runs where perm[j+1] == perm[j] + 1
  Merged m;
  int run_of_src[kMaxModes];
  int runs = 0;
  int run_first[kMaxModes], run_last[kMaxModes];
  for (int j = 0; j < n; ++j) {
    if (j > 0 && perm[j] == perm[j - 1] + 1) {
      run_last[runs - 1] = perm[j];
    } else {
      run_first[runs] = perm[j];
      run_last[runs] = perm[j];
      runs++;
    }
  }
  // merged src mode index = rank of run by its first src mode
  int order[kMaxModes];
  for (int r = 0; r < runs; ++r) order[r] = r;
  std::sort(order, order + runs, [&](int a, int b) { return run_first[a] < run_first[b]; });
  int merged_index_of_run[kMaxModes];
  for (int i = 0; i < runs; ++i) merged_index_of_run[order[i]] = i;
  m.nm = runs;
  for (int r = 0; r < runs; ++r) {
    m.perm[r] = merged_index_of_run[r];
    m.first[merged_index_of_run[r]] = run_first[r];
    m.last[merged_index_of_run[r]] = run_last[r];
  }
  (void)run_of_src;
  return m;
}

hipError_t build_permute(PermuteLaunch* pl, int n, const int* perm, const BlockList& bl,
                         const std::vector<int64_t>* dst_offs, hipStream_t s,
                         const std::vector<uint8_t>* fresh = nullptr) {
  Merged m = merge_modes(n, perm);
  if (m.nm == 0) {  // 0-dim tensor: one mode of extent 1
    m.nm = 1;
    m.perm[0] = 0;
    m.first[0] = 0;
    m.last[0] = -1;
  }
  pl->nm = m.nm;
  for (int i = 0; i < kMaxModes; ++i) pl->perm[i] = i < m.nm ? m.perm[i] : 0;
  pl->kind = (m.nm <= 1 || m.perm[0] == 0) ? 0 : 1;
  int q = 0;
  for (int i = 0; i < m.nm; ++i)
    if (m.perm[i] == 0) q = i;
  std::vector<PermBlock> blocks(bl.nblk());
  std::vector<int64_t> prefix(bl.nblk() + 1, 0);
  for (int64_t b = 0; b < bl.nblk(); ++b) {
    PermBlock& pb = blocks[b];
    std::memset(&pb, 0, sizeof(pb));
    pb.off_src = bl.offs[b];
    pb.off_dst = dst_offs ? (*dst_offs)[b] : bl.offs[b];
    pb.fresh = fresh ? (*fresh)[b] : 0;
    const int32_t* d = bl.d(b);
    int64_t sdim[kMaxModes], sstr[kMaxModes];
    int64_t acc = 1;
    for (int i = 0; i < m.nm; ++i) {  // merged src modes
      int64_t prod = 1;
      for (int k = m.first[i]; k <= m.last[i]; ++k) prod *= d[k];
      sdim[i] = prod;
      sstr[i] = acc;
      acc *= prod;
    }
    pb.nelem = (int32_t)acc;
    for (int i = 0; i < kMaxModes; ++i) {
      if (i < m.nm) {
        pb.ddim[i] = (uint32_t)sdim[m.perm[i]];
        pb.sstr[i] = sstr[m.perm[i]];
      } else {
        pb.ddim[i] = 1;
      }
      fastdiv_consts(std::max<uint32_t>(pb.ddim[i], 1), &pb.dmul[i], &pb.dshr[i]);
    }
    prefix[b + 1] = prefix[b] + permute_block_tiles(m.nm, q, pb.ddim, pl->kind);
  }
  pl->nblk = bl.nblk();
  pl->tiles = prefix.back();
  hipError_t e = pl->blocks.upload(blocks.data(), blocks.size() * sizeof(PermBlock), s);
  if (e != hipSuccess) return e;
  return pl->prefix.upload(prefix.data(), prefix.size() * sizeof(int64_t), s);
}

std::vector<int> seq_concat(const std::vector<int>& a, const std::vector<int>& b) {
  std::vector<int> r(a);
  r.insert(r.end(), b.begin(), b.end());
  return r;
}
bool is_identity(const std::vector<int>& v) {
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] != (int)i) return false;
  return true;
}

int64_t block_size(const BlockList& T, int64_t b) {
  int64_t n = 1;
  for (int i = 0; i < T.N; ++i) n *= T.d(b)[i];
  return n;
}

// Column-major strides of block b of T.
void block_strides(const BlockList& T, int64_t b, int64_t* str) {
  int64_t a = 1;
  for (int i = 0; i < T.N; ++i) {
    str[i] = a;
    a *= T.d(b)[i];
  }
}

int64_t uf_find(std::vector<int64_t>& par, int64_t x) {
  while (par[x] != x) x = par[x] = par[par[x]];
  return x;
}

// Matrix-form FIBER tables (see FiberMVGroup) for the groups comp_granges.
// Returns "" when built, "-" when the form does not apply (then nothing is
// set), or an error message.
std::string build_fiber_mv(const Structure& st, const BlockList& TA, const BlockList& TB,
                           const BlockList& bR, const std::vector<PairSpec>& pairs,
                           const std::vector<std::pair<size_t, size_t>>& granges,
                           const std::vector<std::vector<size_t>>& comp_granges,
                           const std::vector<std::pair<int, int>>& fm,
                           const std::vector<std::pair<int, int>>& gm, const std::vector<int>& kA,
                           const std::vector<int>& kB, bool a2, const double* pair_alpha,
                           hipStream_t s, ExecPlan* ep) {
  const int nf = (int)fm.size(), nk = st.nk;
  auto blkA = [&](const PairSpec& p) { return a2 ? p.blk2 : p.blk1; };
  auto blkB = [&](const PairSpec& p) { return a2 ? p.blk1 : p.blk2; };
  std::vector<FiberMVGroup> groups(comp_granges.size());
  std::vector<FiberSlot> slots;
  std::vector<FiberOutElem> outs;
  std::vector<FiberCoef> coefs;
  std::vector<int32_t> wgg;
  int64_t wgs = 0, total_fib = 0;
  int smax = 1;
  for (size_t c = 0; c < comp_granges.size(); ++c) {
    FiberMVGroup& G = groups[c];
    std::memset(&G, 0, sizeof(G));
    const int64_t a0 = blkA(pairs[granges[comp_granges[c][0]].first]);
    int64_t nfib = 1;
    for (int i = 0; i < kMaxModes; ++i) G.fdim[i] = 1;
    for (int i = 0; i < nf; ++i) {
      G.fdim[i] = TA.d(a0)[fm[(size_t)i].first];
      nfib *= G.fdim[i];
    }
    if (nfib > INT32_MAX) return "-";
    G.nfib = nfib;
    // slots: (A block, k) in first-appearance order of the A blocks
    G.s0 = (int32_t)slots.size();
    std::map<int64_t, int32_t> slot0;  // A block -> its first slot (local)
    std::vector<int64_t> sA(kMaxModes);
    for (size_t gi : comp_granges[c])
      for (size_t q = granges[gi].first; q < granges[gi].second; ++q) {
        const int64_t ba = blkA(pairs[q]);
        if (slot0.count(ba)) continue;
        for (int i = 0; i < nf; ++i)
          if (TA.d(ba)[fm[(size_t)i].first] != G.fdim[i]) return "-";  // not one sector
        block_strides(TA, ba, sA.data());
        slot0[ba] = (int32_t)(slots.size() - G.s0);
        int64_t K = 1;
        for (int i = 0; i < nk; ++i) K *= TA.d(ba)[kA[(size_t)i]];
        std::vector<int32_t> kc((size_t)std::max(nk, 1), 0);
        for (int64_t k = 0; k < K; ++k) {
          FiberSlot sl;
          std::memset(&sl, 0, sizeof(sl));
          sl.off_a = TA.offs[ba];
          for (int i = 0; i < nf; ++i) sl.astr[i] = (int32_t)sA[(size_t)fm[(size_t)i].first];
          int64_t oa = 0;
          for (int i = 0; i < nk; ++i) oa += (int64_t)kc[(size_t)i] * sA[(size_t)kA[(size_t)i]];
          sl.ka = (int32_t)oa;
          slots.push_back(sl);
          for (int i = 0; i < nk; ++i) {
            if (++kc[(size_t)i] < TA.d(ba)[kA[(size_t)i]]) break;
            kc[(size_t)i] = 0;
          }
        }
      }
    G.nslot = (int32_t)(slots.size() - G.s0);
    // outputs: (output block, g) in plan order; coefficients per pair
    G.o0 = (int32_t)outs.size();
    G.w0 = (int32_t)coefs.size();
    std::map<std::pair<int32_t, int32_t>, int> seen;
    for (size_t gi : comp_granges[c]) {
      const size_t q0 = granges[gi].first, q1 = granges[gi].second;
      const int64_t br = pairs[q0].blkR;
      int64_t sR[kMaxModes];
      block_strides(bR, br, sR);
      int64_t ng = 1;
      for (const auto& m : gm) ng *= bR.d(br)[m.second];
      const int32_t out0 = (int32_t)(outs.size() - G.o0);
      {
        std::vector<int32_t> cc(gm.size(), 0);
        for (int64_t g = 0; g < ng; ++g) {
          FiberOutElem oe;
          std::memset(&oe, 0, sizeof(oe));
          oe.off_c = bR.offs[br];
          for (int i = 0; i < nf; ++i) oe.cstr[i] = (int32_t)sR[fm[(size_t)i].second];
          int64_t o = 0;
          for (size_t j = 0; j < gm.size(); ++j) o += (int64_t)cc[j] * sR[gm[j].second];
          oe.go = (int32_t)o;
          outs.push_back(oe);
          for (size_t j = 0; j < gm.size(); ++j) {
            if (++cc[j] < bR.d(br)[gm[j].second]) break;
            cc[j] = 0;
          }
        }
      }
      for (size_t q = q0; q < q1; ++q) {
        const PairSpec& P = pairs[q];
        const int64_t ba = blkA(P), bb = blkB(P);
        int64_t sB[kMaxModes];
        block_strides(TB, bb, sB);
        const double ar = pair_alpha ? pair_alpha[2 * q] : 1.0;
        const double ai = pair_alpha ? pair_alpha[2 * q + 1] : 0.0;
        int64_t K = 1;
        for (int i = 0; i < nk; ++i) K *= TA.d(ba)[kA[(size_t)i]];
        std::vector<int32_t> kc((size_t)std::max(nk, 1), 0);
        for (int64_t k = 0; k < K; ++k) {
          int64_t ob = 0;
          for (int i = 0; i < nk; ++i) ob += (int64_t)kc[(size_t)i] * sB[kB[(size_t)i]];
          std::vector<int32_t> cc(gm.size(), 0);
          for (int64_t g = 0; g < ng; ++g) {
            int64_t og = 0;
            for (size_t j = 0; j < gm.size(); ++j) og += (int64_t)cc[j] * sB[gm[j].first];
            FiberCoef cf;
            std::memset(&cf, 0, sizeof(cf));
            cf.off_b = TB.offs[bb] + ob + og;
            cf.alpha_re = ar;
            cf.alpha_im = ai;
            cf.slot = slot0[ba] + (int32_t)k;
            cf.out = out0 + (int32_t)g;
            if (seen[{cf.slot, cf.out}]++) return "-";  // one pair per (slot, output)
            if (cf.slot >= kFiberSlots) return "-";
            outs[(size_t)(G.o0 + cf.out)].mask |= 1u << cf.slot;
            coefs.push_back(cf);
            for (size_t j = 0; j < gm.size(); ++j) {
              if (++cc[j] < bR.d(br)[gm[j].second]) break;
              cc[j] = 0;
            }
          }
          for (int i = 0; i < nk; ++i) {
            if (++kc[(size_t)i] < TA.d(ba)[kA[(size_t)i]]) break;
            kc[(size_t)i] = 0;
          }
        }
      }
    }
    G.nout = (int32_t)(outs.size() - G.o0);
    G.nw = (int32_t)(coefs.size() - G.w0);
    if (G.nslot > kFiberSlots || G.nout > kFiberOuts || G.nslot < 1) return "-";
    smax = std::max(smax, (int)G.nslot);
    G.wg0 = wgs;
    const int64_t n = (nfib + kFibersPerWGMV - 1) / kFibersPerWGMV;
    for (int64_t w = 0; w < n; ++w) wgg.push_back((int32_t)c);
    wgs += n;
    total_fib += nfib;
  }
  if (total_fib < 256) return "-";
  if (coefs.empty()) coefs.push_back(FiberCoef{});
  if (wgg.empty()) wgg.push_back(0);
  ep->f_a_is_2 = a2;
  ep->f_nf = nf;
  ep->f_wgs = wgs;
  ep->f_mv = true;
  ep->f_smax = smax <= 4 ? 4 : (smax <= 8 ? 8 : kFiberSlots);
  ep->hfm_groups = groups;
  ep->hfm_slots = slots;
  ep->hfm_outs = outs;
  ep->hfm_coefs = coefs;
  ep->f_fm = fm;
  hipError_t e;
  if ((e = ep->fm_groups.upload(groups.data(), groups.size() * sizeof(FiberMVGroup), s)) !=
          hipSuccess ||
      (e = ep->fm_slots.upload(slots.data(), slots.size() * sizeof(FiberSlot), s)) !=
          hipSuccess ||
      (e = ep->fm_outs.upload(outs.data(), outs.size() * sizeof(FiberOutElem), s)) !=
          hipSuccess ||
      (e = ep->fm_coefs.upload(coefs.data(), coefs.size() * sizeof(FiberCoef), s)) !=
          hipSuccess ||
      (e = ep->f_wgg.upload(wgg.data(), wgg.size() * sizeof(int32_t), s)) != hipSuccess)
    return std::string("device table upload failed: ") + hipGetErrorString(e);
  return "";
}

// FIBER tables. Returns false (nothing built) when the shape does not suit:
// the smaller operand is not at most a quarter of the larger, the larger has
// no free mode, too few fibers, or too many B-free elements per output block.
// On an upload failure returns true with *err set.
bool build_fiber_tables(const Structure& st, const BlockList& b1, const BlockList& b2,
                        const BlockList& bR, const std::vector<PairSpec>& pairs,
                        const std::vector<std::pair<size_t, size_t>>& granges,
                        const double* pair_alpha, hipStream_t s, ExecPlan* ep,
                        std::string* err) {
  const int NR = st.NR, nk = st.nk;
  // operand sizes over the blocks the pairs touch
  int64_t size1 = 0, size2 = 0;
  {
    std::vector<char> u1(b1.nblk(), 0), u2(b2.nblk(), 0);
    for (const PairSpec& p : pairs) {
      if (!u1[p.blk1]) u1[p.blk1] = 1, size1 += block_size(b1, p.blk1);
      if (!u2[p.blk2]) u2[p.blk2] = 1, size2 += block_size(b2, p.blk2);
    }
  }
  const bool a2 = size2 > size1;
  const BlockList& TA = a2 ? b2 : b1;
  const BlockList& TB = a2 ? b1 : b2;
  if (4 * (a2 ? size1 : size2) > (a2 ? size2 : size1)) return false;
  const int srcA = a2 ? 2 : 1;
  // A-free modes in A's mode order: (A position, R position)
  std::vector<std::pair<int, int>> fm;
  std::vector<std::pair<int, int>> gm;  // B-free modes in R order: (B position, R position)
  for (int r = 0; r < NR; ++r) {
    if (st.r_src[r] == srcA)
      fm.push_back({st.r_pos[r], r});
    else
      gm.push_back({st.r_pos[r], r});
  }
  std::sort(fm.begin(), fm.end());
  const int nf = (int)fm.size();
  if (nf == 0 || nf > kMaxModes) return false;
  std::vector<int> kA(nk), kB(nk);
  for (int i = 0; i < nk; ++i) {
    kA[i] = a2 ? st.k2[i] : st.k1[i];
    kB[i] = a2 ? st.k1[i] : st.k2[i];
  }
  auto blkA = [&](const PairSpec& p) { return a2 ? p.blk2 : p.blk1; };
  auto blkB = [&](const PairSpec& p) { return a2 ? p.blk1 : p.blk2; };

  // groups: connected components of (A block) -- (output block) over pairs
  const int64_t NA = TA.nblk();
  std::vector<int64_t> par((size_t)(NA + bR.nblk()));
  std::iota(par.begin(), par.end(), 0);
  for (const PairSpec& p : pairs) {
    const int64_t x = uf_find(par, blkA(p)), y = uf_find(par, NA + p.blkR);
    if (x != y) par[x] = y;
  }
  std::map<int64_t, int> comp_of_root;
  std::vector<std::vector<size_t>> comp_granges;  // indices into granges, plan order
  for (size_t gi = 0; gi < granges.size(); ++gi) {
    const int64_t root = uf_find(par, NA + pairs[granges[gi].first].blkR);
    auto it = comp_of_root.find(root);
    int c;
    if (it == comp_of_root.end()) {
      c = (int)comp_granges.size();
      comp_of_root[root] = c;
      comp_granges.emplace_back();
    } else {
      c = it->second;
    }
    comp_granges[(size_t)c].push_back(gi);
  }

  // The matrix form when every group has few enough slots and outputs.
  {
    bool fits = true;
    for (size_t c = 0; c < comp_granges.size() && fits; ++c) {
      std::map<int64_t, int64_t> kOfA;  // distinct A blocks -> K
      int64_t nout = 0;
      for (size_t gi : comp_granges[c]) {
        const size_t q0 = granges[gi].first, q1 = granges[gi].second;
        int64_t ng = 1;
        for (const auto& m : gm) ng *= bR.d(pairs[q0].blkR)[m.second];
        nout += ng;
        for (size_t q = q0; q < q1; ++q) {
          int64_t K = 1;
          for (int i = 0; i < nk; ++i) K *= TA.d(blkA(pairs[q]))[kA[(size_t)i]];
          kOfA[blkA(pairs[q])] = K;
        }
      }
      int64_t nslot = 0;
      for (const auto& kv : kOfA) nslot += kv.second;
      fits = nslot >= 1 && nslot <= kFiberSlots && nout <= kFiberOuts;
    }
    if (fits) {
      std::string r = build_fiber_mv(st, TA, TB, bR, pairs, granges, comp_granges, fm, gm, kA, kB,
                                     a2, pair_alpha, s, ep);
      if (r != "-") {
        *err = r;
        return true;
      }
    }
  }

  std::vector<FiberGroup> fg(comp_granges.size());
  std::vector<FiberOut> fo;
  std::vector<FiberPair> fp(pairs.size());
  std::vector<int32_t> koffs, gofs_c, gofs_b;
  int64_t total_fib = 0, wgs = 0;
  std::vector<int32_t> wgg;
  for (size_t c = 0; c < comp_granges.size(); ++c) {
    FiberGroup& G = fg[c];
    std::memset(&G, 0, sizeof(G));
    const int64_t a0 = blkA(pairs[granges[comp_granges[c][0]].first]);
    int64_t nfib = 1;
    for (int i = 0; i < kMaxModes; ++i) G.fdim[i] = 1;
    for (int i = 0; i < nf; ++i) {
      G.fdim[i] = TA.d(a0)[fm[(size_t)i].first];
      nfib *= G.fdim[i];
    }
    if (nfib > INT32_MAX) return false;
    G.nfib = nfib;
    G.r0 = (int32_t)fo.size();
    for (size_t gi : comp_granges[c]) {
      const size_t q0 = granges[gi].first, q1 = granges[gi].second;
      const int64_t br = pairs[q0].blkR;
      int64_t sR[kMaxModes];
      block_strides(bR, br, sR);
      FiberOut O;
      std::memset(&O, 0, sizeof(O));
      O.off_c = bR.offs[br];
      for (int i = 0; i < nf; ++i) O.cstr[i] = (int32_t)sR[fm[(size_t)i].second];
      // B-free elements of this output block, column-major over gm (R order)
      int64_t ng = 1;
      for (const auto& m : gm) ng *= bR.d(br)[m.second];
      if (ng > 1024) return false;
      O.g0 = (int32_t)gofs_c.size();
      O.ng = (int32_t)ng;
      {
        std::vector<int32_t> cc(gm.size(), 0);
        for (int64_t g = 0; g < ng; ++g) {
          int64_t o = 0;
          for (size_t j = 0; j < gm.size(); ++j) o += (int64_t)cc[j] * sR[gm[j].second];
          gofs_c.push_back((int32_t)o);
          for (size_t j = 0; j < gm.size(); ++j) {
            if (++cc[j] < bR.d(br)[gm[j].second]) break;
            cc[j] = 0;
          }
        }
      }
      O.p0 = (int32_t)q0;
      O.p1 = (int32_t)q1;
      for (size_t q = q0; q < q1; ++q) {
        const PairSpec& P = pairs[q];
        const int64_t ba = blkA(P), bb = blkB(P);
        for (int i = 0; i < nf; ++i)
          if (TA.d(ba)[fm[(size_t)i].first] != G.fdim[i]) return false;  // not one sector
        int64_t sA[kMaxModes], sB[kMaxModes];
        block_strides(TA, ba, sA);
        block_strides(TB, bb, sB);
        FiberPair& X = fp[q];
        std::memset(&X, 0, sizeof(X));
        X.off_a = TA.offs[ba];
        X.off_b = TB.offs[bb];
        if (pair_alpha) {
          X.alpha_re = pair_alpha[2 * q];
          X.alpha_im = pair_alpha[2 * q + 1];
        } else {
          X.alpha_re = 1.0;
        }
        for (int i = 0; i < nf; ++i) X.astr[i] = (int32_t)sA[fm[(size_t)i].first];
        // contracted multi-index, first contracted mode fastest
        int64_t K = 1;
        for (int i = 0; i < nk; ++i) K *= TA.d(ba)[kA[(size_t)i]];
        X.kofs = (int32_t)(koffs.size() / 2);
        X.nk = (int32_t)K;
        {
          std::vector<int32_t> kc((size_t)std::max(nk, 1), 0);
          for (int64_t k = 0; k < K; ++k) {
            int64_t oa = 0, ob = 0;
            for (int i = 0; i < nk; ++i) {
              oa += (int64_t)kc[(size_t)i] * sA[kA[(size_t)i]];
              ob += (int64_t)kc[(size_t)i] * sB[kB[(size_t)i]];
            }
            koffs.push_back((int32_t)oa);
            koffs.push_back((int32_t)ob);
            for (int i = 0; i < nk; ++i) {
              if (++kc[(size_t)i] < TA.d(ba)[kA[(size_t)i]]) break;
              kc[(size_t)i] = 0;
            }
          }
        }
        X.gb0 = (int32_t)gofs_b.size();
        {
          std::vector<int32_t> cc(gm.size(), 0);
          for (int64_t g = 0; g < ng; ++g) {
            int64_t o = 0;
            for (size_t j = 0; j < gm.size(); ++j) o += (int64_t)cc[j] * sB[gm[j].first];
            gofs_b.push_back((int32_t)o);
            for (size_t j = 0; j < gm.size(); ++j) {
              if (++cc[j] < bR.d(br)[gm[j].second]) break;
              cc[j] = 0;
            }
          }
        }
      }
      fo.push_back(O);
    }
    G.r1 = (int32_t)fo.size();
    G.wg0 = wgs;
    const int64_t n = (nfib + kFibersPerWG - 1) / kFibersPerWG;
    for (int64_t w = 0; w < n; ++w) wgg.push_back((int32_t)c);
    wgs += n;
    total_fib += nfib;
  }
  if (total_fib < 256) return false;
  if (koffs.empty()) koffs.push_back(0), koffs.push_back(0);
  if (gofs_c.empty()) gofs_c.push_back(0);
  if (gofs_b.empty()) gofs_b.push_back(0);
  if (wgg.empty()) wgg.push_back(0);
  ep->f_a_is_2 = a2;
  ep->f_nf = nf;
  ep->f_wgs = wgs;
  hipError_t e;
  if ((e = ep->f_groups.upload(fg.data(), fg.size() * sizeof(FiberGroup), s)) != hipSuccess ||
      (e = ep->f_outs.upload(fo.data(), fo.size() * sizeof(FiberOut), s)) != hipSuccess ||
      (e = ep->f_pairs.upload(fp.data(), fp.size() * sizeof(FiberPair), s)) != hipSuccess ||
      (e = ep->f_koffs.upload(koffs.data(), koffs.size() * sizeof(int32_t), s)) != hipSuccess ||
      (e = ep->f_gofs_c.upload(gofs_c.data(), gofs_c.size() * sizeof(int32_t), s)) !=
          hipSuccess ||
      (e = ep->f_gofs_b.upload(gofs_b.data(), gofs_b.size() * sizeof(int32_t), s)) !=
          hipSuccess ||
      (e = ep->f_wgg.upload(wgg.data(), wgg.size() * sizeof(int32_t), s)) != hipSuccess)
    *err = std::string("device table upload failed: ") + hipGetErrorString(e);
  return true;
}

}  // namespace

hipError_t build_permute_launch(PermuteLaunch* pl, int n, const int* perm, const BlockList& bl,
                                const std::vector<int64_t>* dst_offs, hipStream_t s,
                                const std::vector<uint8_t>* fresh) {
  return build_permute(pl, n, perm, bl, dst_offs, s, fresh);
}

// ----------------------------------------------------------- exec plan
// Split-K tables of a GEMM plan (FP64). A plan whose output blocks make few
// 64 x 64 tiles gives each workgroup a whole tile's contracted extent -- an
// inner product, or the cotangent of a small tensor such as an MPO factor,
// is one tile over millions of K values, i.e. one workgroup. Such a plan
// also gets a split form: the pairs of each output block are laid end to end
// along K, that range is cut into S slices of at least kKsSlice values, and
// every (tile, slice) is a workgroup writing a partial block; the reduce
// kernel adds the partials in slice order. Nothing is built (ks_grid = 0)
// for plans with enough tiles, or when no block would be split.
std::string build_ksplit(ExecPlan* ep, hipStream_t s) {
  constexpr int64_t kKsTiles = 128;       // split only below this many tiles
  constexpr int64_t kKsSlice = 2048;      // contracted values per slice, at least
  constexpr int64_t kKsTarget = 2048;     // workgroups aimed for
  constexpr int64_t kKsPartial = 1 << 21; // partial elements, at most (32 MiB complex)
  const size_t ng = ep->h_groups.size();
  if (ng == 0) return "";
  std::vector<int64_t> tiles(ng), klen(ng), S(ng, 1);
  int64_t ttot = 0;
  for (size_t g = 0; g < ng; ++g) {
    const GemmGroup& G = ep->h_groups[g];
    tiles[g] = (int64_t)((G.M + 63) / 64) * ((G.N + 63) / 64);
    klen[g] = 0;
    for (int p = G.p0; p < G.p1; ++p) klen[g] += std::max(ep->h_pairs[p].K, 0);
    ttot += tiles[g];
  }
  if (ttot == 0 || ttot >= kKsTiles) return "";
  // dot-like: every block <= kKsDotMN elements; skinny: every block has a
  // side <= kKsSkinny (matrix times vectors). The skinny kernel gives each
  // workgroup 256 rows of the long side, so its slices are cut finer
  // (>= 64 values) to reach the same number of workgroups.
  bool dot = true, skinny = true;
  for (const GemmGroup& G : ep->h_groups) {
    if ((int64_t)G.M * G.N > kKsDotMN) dot = false;
    if (std::min(G.M, G.N) > kKsSkinny) skinny = false;
  }
  skinny = skinny && !dot;
  // skinny: are the long side's rows contiguous along K? A(m, k) is when
  // ta, B(n, k) when !tb; the first block decides which operand holds the
  // long side (a plan whose blocks disagree is read correctly either way,
  // only less coalesced). K-contiguous rows go 16 to a workgroup with lanes
  // along K and slices of >= 512 values; the others 256 to a workgroup, one
  // row per thread, slices of >= 64.
  const bool kc = skinny && (ep->h_groups[0].N <= kKsSkinny ? ep->ta != 0 : ep->tb == 0);
  const int rows_wg = kKsSkinnyRows[kc ? 1 : 0];
  int64_t chunks = 0;
  for (const GemmGroup& G : ep->h_groups) chunks += (std::max(G.M, G.N) + rows_wg - 1) / rows_wg;
  const int64_t smax = std::max<int64_t>(1, kKsTarget / (skinny ? chunks : ttot));
  const int64_t kslice = skinny ? (kc ? 512 : 64) : kKsSlice;
  int64_t partial = 0;
  for (size_t g = 0; g < ng; ++g) {
    S[g] = std::max<int64_t>(1, std::min<int64_t>(klen[g] / kslice, smax));
    partial += S[g] * ep->h_groups[g].M * ep->h_groups[g].N;
  }
  while (partial > kKsPartial) {
    partial = 0;
    for (size_t g = 0; g < ng; ++g) {
      S[g] = std::max<int64_t>(1, S[g] / 2);
      partial += S[g] * ep->h_groups[g].M * ep->h_groups[g].N;
    }
  }
  if (*std::max_element(S.begin(), S.end()) <= 1) return "";
  // sub-pairs: a pair cut at slice boundaries keeps its leading dimensions;
  // its operands start k0 values further along K
  std::vector<GemmPair> sp;
  std::vector<GemmGroup> sg;
  std::vector<KsReduce> red(ng);
  int64_t poff = 0, e0 = 0, maxk = 0;
  for (size_t g = 0; g < ng; ++g) {
    const GemmGroup& G = ep->h_groups[g];
    const int64_t mn = (int64_t)G.M * G.N;
    red[g] = KsReduce{poff, G.off_c, e0, G.M, G.N, G.ldc, (int32_t)S[g]};
    maxk = std::max(maxk, klen[g]);
    int p = G.p0;
    int64_t kin = 0;  // values of pair p already given to earlier slices
    for (int64_t q = 0; q < S[g]; ++q) {
      const int64_t want = klen[g] * (q + 1) / S[g] - klen[g] * q / S[g];
      GemmGroup h = G;
      h.off_c = poff + q * mn;
      h.ldc = std::max(G.M, 1);
      h.p0 = (int32_t)sp.size();
      int64_t got = 0;
      while (got < want && p < G.p1) {
        const GemmPair& pr = ep->h_pairs[p];
        const int64_t left = std::max(pr.K, 0) - kin;
        if (left <= 0) {
          ++p;
          kin = 0;
          continue;
        }
        const int64_t take = std::min(left, want - got);
        GemmPair x = pr;
        x.K = (int32_t)take;
        x.off_a += ep->ta ? kin : kin * pr.lda;
        x.off_b += ep->tb ? kin * pr.ldb : kin;
        sp.push_back(x);
        got += take;
        kin += take;
      }
      h.p1 = (int32_t)sp.size();
      sg.push_back(h);
    }
    poff += S[g] * mn;
    e0 += mn;
  }
  ExecPlan tab;
  tab.h_groups = sg;
  tab.h_pairs = sp;
  std::vector<int2> tiles_tab = xcd_tile_table(tab, gemm_cfg(0));
  // skinny: (entry, first row) per workgroup, 256 rows of the long side
  std::vector<int2> rows;
  if (skinny)
    for (size_t j = 0; j < sg.size(); ++j) {
      const int R = sg[j].N <= kKsSkinny ? sg[j].M : sg[j].N;
      for (int r0 = 0; r0 < R; r0 += rows_wg) rows.push_back(int2{(int)j, r0});
    }
  hipError_t e;
  if ((e = ep->ks_pairs.upload(sp.data(), std::max<size_t>(sp.size(), 1) * sizeof(GemmPair), s)) !=
          hipSuccess ||
      (e = ep->ks_groups.upload(sg.data(), sg.size() * sizeof(GemmGroup), s)) != hipSuccess ||
      (e = ep->ks_tiles.upload(tiles_tab.data(), tiles_tab.size() * sizeof(int2), s)) !=
          hipSuccess ||
      (e = ep->ks_red.upload(red.data(), red.size() * sizeof(KsReduce), s)) != hipSuccess ||
      (skinny && (e = ep->ks_rows.upload(rows.data(), rows.size() * sizeof(int2), s)) !=
                     hipSuccess))
    return std::string("split-K table upload failed: ") + hipGetErrorString(e);
  ep->ks_grid = (int64_t)tiles_tab.size();
  ep->ks_elems = poff;
  ep->ks_total = e0;
  ep->ks_maxk = maxk;
  ep->ks_nred = (int)ng;
  ep->ks_maxs = (int)*std::max_element(S.begin(), S.end());
  ep->ks_ngroups = (int64_t)sg.size();
  ep->ks_dot = dot;
  ep->ks_dot1 = dot;
  for (const GemmGroup& G : ep->h_groups)
    if ((int64_t)G.M * G.N != 1) ep->ks_dot1 = false;
  ep->ks_skinny = skinny;
  ep->ks_kc = kc;
  ep->ks_nrows = (int64_t)rows.size();
  return "";
}

std::string build_exec_plan(const Structure& st, const BlockList& b1, const BlockList& b2,
                            const BlockList& bR, const std::vector<PairSpec>& pairs,
                            const double* pair_alpha, hipStream_t s, ExecPlan* ep,
                            int backend) {
  ep->st = st;
  const int NR = st.NR, nk = st.nk;
  // per-pair shape figures
  double mnk = 0.0, mn = 0.0, flops = 0.0;
  int64_t maxdim = 0;
  for (size_t p = 0; p < pairs.size(); ++p) {
    const int32_t* d1 = b1.d(pairs[p].blk1);
    const int32_t* d2 = b2.d(pairs[p].blk2);
    double K = 1, M = 1, N = 1;
    for (int i = 0; i < nk; ++i) K *= d1[st.k1[i]];
    for (int r = 0; r < NR; ++r) {
      if (st.r_src[r] == 1)
        M *= d1[st.r_pos[r]];
      else
        N *= d2[st.r_pos[r]];
    }
    mnk += M * N * K;
    mn += M * N;
    flops += 2.0 * M * N * K;
    maxdim = std::max<int64_t>(maxdim, (int64_t)std::max(M, N));
  }
  ep->flops = flops;
  ep->max_dim = maxdim;
  const double kavg = mn > 0 ? mnk / mn : 0.0;
  const bool use_gemm = backend == kBackendGeneric ? false
                       : backend == kBackendGemm  ? nk > 0
                                                  : nk > 0 && kavg >= 24.0;

  // groups: consecutive pairs sharing blkR
  std::vector<std::pair<size_t, size_t>> granges;
  for (size_t p = 0; p < pairs.size();) {
    size_t q = p;
    while (q < pairs.size() && pairs[q].blkR == pairs[p].blkR) ++q;
    granges.push_back({p, q});
    p = q;
  }
  auto alpha_of = [&](size_t p, double* re, double* im) {
    if (pair_alpha) {
      *re = pair_alpha[2 * p];
      *im = pair_alpha[2 * p + 1];
    } else {
      *re = 1.0;
      *im = 0.0;
    }
  };

  if (!use_gemm) {
    std::string ferr;
    if (backend != kBackendGeneric &&
        build_fiber_tables(st, b1, b2, bR, pairs, granges, pair_alpha, s, ep, &ferr)) {
      ep->strategy = ExecPlan::FIBER;
      return ferr;
    }
    ep->strategy = ExecPlan::DIRECT;
    std::vector<DirectPair> dp(pairs.size());
    std::vector<DirectGroup> dg(granges.size());
    std::vector<int32_t> koffs;
    std::vector<int64_t> prefix(granges.size() + 1, 0);
    for (size_t p = 0; p < pairs.size(); ++p) {
      const int32_t* d1 = b1.d(pairs[p].blk1);
      const int32_t* d2 = b2.d(pairs[p].blk2);
      int64_t s1[kMaxModes], s2[kMaxModes];
      {
        int64_t a = 1;
        for (int i = 0; i < st.N1; ++i) {
          s1[i] = a;
          a *= d1[i];
        }
        int64_t b = 1;
        for (int i = 0; i < st.N2; ++i) {
          s2[i] = b;
          b *= d2[i];
        }
      }
      DirectPair& x = dp[p];
      std::memset(&x, 0, sizeof(x));
      x.off_a = b1.offs[pairs[p].blk1];
      x.off_b = b2.offs[pairs[p].blk2];
      alpha_of(p, &x.alpha_re, &x.alpha_im);
      for (int r = 0; r < NR; ++r) {
        if (st.r_src[r] == 1)
          x.cstr_a[r] = (int32_t)s1[st.r_pos[r]];
        else
          x.cstr_b[r] = (int32_t)s2[st.r_pos[r]];
      }
      int64_t K = 1;
      for (int i = 0; i < nk; ++i) K *= d1[st.k1[i]];
      x.kofs = (int32_t)(koffs.size() / 2);
      x.nk = (int32_t)K;
      // enumerate contracted multi-index, first contracted mode fastest
      int64_t c[kMaxModes] = {0};
      for (int64_t k = 0; k < K; ++k) {
        int64_t oa = 0, ob = 0;
        for (int i = 0; i < nk; ++i) {
          oa += c[i] * s1[st.k1[i]];
          ob += c[i] * s2[st.k2[i]];
        }
        koffs.push_back((int32_t)oa);
        koffs.push_back((int32_t)ob);
        for (int i = 0; i < nk; ++i) {
          if (++c[i] < d1[st.k1[i]]) break;
          c[i] = 0;
        }
      }
    }
    for (size_t g = 0; g < granges.size(); ++g) {
      DirectGroup& y = dg[g];
      std::memset(&y, 0, sizeof(y));
      const int64_t br = pairs[granges[g].first].blkR;
      y.off_c = bR.offs[br];
      int64_t ne = 1;
      for (int r = 0; r < kMaxModes; ++r) y.cdim[r] = 1;
      for (int r = 0; r < NR; ++r) {
        y.cdim[r] = bR.d(br)[r];
        ne *= y.cdim[r];
      }
      y.nelem = (int32_t)ne;
      y.p0 = (int32_t)granges[g].first;
      y.p1 = (int32_t)granges[g].second;
      y.wg0 = prefix[g];
      prefix[g + 1] = prefix[g] + direct_group_tiles(&y, NR);
    }
    if (koffs.empty()) koffs.push_back(0), koffs.push_back(0);
    ep->d_ngroups = (int64_t)granges.size();
    ep->d_tiles = prefix.back();
    // workgroup -> group
    std::vector<int32_t> wgg((size_t)std::max<int64_t>(ep->d_tiles, 1), 0);
    for (size_t g = 0; g < granges.size(); ++g)
      for (int64_t w = prefix[g]; w < prefix[g + 1]; ++w) wgg[(size_t)w] = (int32_t)g;
    hipError_t e;
    if ((e = ep->d_pairs.upload(dp.data(), dp.size() * sizeof(DirectPair), s)) != hipSuccess ||
        (e = ep->d_groups.upload(dg.data(), dg.size() * sizeof(DirectGroup), s)) != hipSuccess ||
        (e = ep->d_koffs.upload(koffs.data(), koffs.size() * sizeof(int32_t), s)) != hipSuccess ||
        (e = ep->d_wgg.upload(wgg.data(), wgg.size() * sizeof(int32_t), s)) != hipSuccess)
      return std::string("device table upload failed: ") + hipGetErrorString(e);
    return "";
  }

  // ------------------------------------------------------------- GEMM
  ep->strategy = ExecPlan::GEMM;
  // orientation
  std::vector<int> m1, n2;  // R-order positions in T1 / T2 of free modes
  bool t1_first = true, t2_first = true;
  {
    bool seen2 = false, seen1 = false;
    for (int r = 0; r < NR; ++r) {
      if (st.r_src[r] == 1) {
        m1.push_back(st.r_pos[r]);
        if (seen2) t1_first = false;
        seen1 = true;
      } else {
        n2.push_back(st.r_pos[r]);
        if (seen1) t2_first = false;
        seen2 = true;
      }
    }
  }
  bool swap = false, permC = false;
  if (t1_first)
    swap = false;
  else if (t2_first)
    swap = true;
  else
    permC = true;
  // P = A operand (rows), Q = B operand (cols)
  const BlockList& P = swap ? b2 : b1;
  const BlockList& Q = swap ? b1 : b2;
  const int NP = swap ? st.N2 : st.N1, NQ = swap ? st.N1 : st.N2;
  std::vector<int> Mp = swap ? n2 : m1;  // rows: positions in P, R order
  std::vector<int> Nq = swap ? m1 : n2;  // cols: positions in Q, R order
  std::vector<int> kP(nk), kQ(nk);
  for (int i = 0; i < nk; ++i) {
    kP[i] = swap ? st.k2[i] : st.k1[i];
    kQ[i] = swap ? st.k1[i] : st.k2[i];
  }
  // candidate X orders: as in P (ascending P positions) or as in Q
  struct Choice {
    std::vector<int> Xp, Xq;
    bool permA, permB;
    int ta, tb;
    int64_t cost;
  };
  auto evaluate = [&](bool by_p) {
    Choice c;
    std::vector<int> idx(nk);
    std::iota(idx.begin(), idx.end(), 0);
    if (by_p)
      std::sort(idx.begin(), idx.end(), [&](int a, int b) { return kP[a] < kP[b]; });
    else
      std::sort(idx.begin(), idx.end(), [&](int a, int b) { return kQ[a] < kQ[b]; });
    for (int i : idx) {
      c.Xp.push_back(kP[i]);
      c.Xq.push_back(kQ[i]);
    }
    c.permA = false;
    c.ta = 0;
    if (is_identity(seq_concat(Mp, c.Xp)))
      c.ta = 0;
    else if (is_identity(seq_concat(c.Xp, Mp)))
      c.ta = 1;
    else
      c.permA = true;
    c.permB = false;
    c.tb = 0;
    if (is_identity(seq_concat(c.Xq, Nq)))
      c.tb = 0;
    else if (is_identity(seq_concat(Nq, c.Xq)))
      c.tb = 1;
    else
      c.permB = true;
    c.cost = (c.permA ? P.nnz : 0) + (c.permB ? Q.nnz : 0);
    return c;
  };
  Choice c1 = evaluate(true), c2 = evaluate(false);
  Choice& ch = (c2.cost < c1.cost) ? c2 : c1;
  ep->swap = swap;
  ep->permC = permC;
  ep->permA = ch.permA;
  ep->permB = ch.permB;
  ep->ta = ch.permA ? 0 : ch.ta;
  ep->tb = ch.permB ? 0 : ch.tb;
  (void)NP;
  (void)NQ;
  // Workspace copies (permuted operands, the pre-permutation C) are packed:
  // block b at the running sum of the sizes before it, whatever the source
  // offsets are (they may be far apart when a batch spans allocations).
  auto packed = [](const BlockList& T, std::vector<int64_t>* offs) {
    offs->resize(T.nblk());
    int64_t acc = 0;
    for (int64_t b = 0; b < T.nblk(); ++b) {
      (*offs)[b] = acc;
      int64_t n = 1;
      for (int i = 0; i < T.N; ++i) n *= T.d(b)[i];
      acc += n;
    }
    return acc;
  };
  std::vector<int64_t> offA, offB, offC;
  ep->nnzA = packed(P, &offA);
  ep->nnzB = packed(Q, &offB);
  ep->nnzC = packed(bR, &offC);
  hipError_t e;
  if (ch.permA) {
    std::vector<int> perm = seq_concat(Mp, ch.Xp);  // new mode i = old mode perm[i]
    if ((e = build_permute(&ep->pA, (int)perm.size(), perm.data(), P, &offA, s)) != hipSuccess)
      return std::string("device table upload failed: ") + hipGetErrorString(e);
  }
  if (ch.permB) {
    std::vector<int> perm = seq_concat(ch.Xq, Nq);
    if ((e = build_permute(&ep->pB, (int)perm.size(), perm.data(), Q, &offB, s)) != hipSuccess)
      return std::string("device table upload failed: ") + hipGetErrorString(e);
  }
  if (permC) {
    // GEMM writes C in (T1 free modes in R order, T2 free modes in R order);
    // R mode r = temp mode perm[r].
    std::vector<int> perm(NR);
    int a = 0, b = (int)m1.size();
    for (int r = 0; r < NR; ++r) perm[r] = (st.r_src[r] == 1) ? a++ : b++;
    // dims of the temp blocks
    BlockList tmp;
    tmp.N = NR;
    tmp.offs = offC;
    tmp.nnz = ep->nnzC;
    tmp.dims.resize(bR.dims.size());
    for (int64_t blk = 0; blk < bR.nblk(); ++blk) {
      const int32_t* d = bR.d(blk);
      int32_t* t = tmp.dims.data() + blk * NR;
      for (int r = 0; r < NR; ++r) t[perm[r]] = d[r];
    }
    if ((e = build_permute(&ep->pC, NR, perm.data(), tmp, &bR.offs, s)) != hipSuccess)
      return std::string("device table upload failed: ") + hipGetErrorString(e);
  }
  // GEMM pairs and groups
  ep->h_pairs.resize(pairs.size());
  ep->h_groups.resize(granges.size());
  ep->group_cost.resize(granges.size());
  for (size_t p = 0; p < pairs.size(); ++p) {
    const int64_t bp = swap ? pairs[p].blk2 : pairs[p].blk1;
    const int64_t bq = swap ? pairs[p].blk1 : pairs[p].blk2;
    const int32_t* dP = P.d(bp);
    const int32_t* dQ = Q.d(bq);
    int64_t M = 1, N = 1, K = 1;
    for (int x : Mp) M *= dP[x];
    for (int x : Nq) N *= dQ[x];
    for (int x : ch.Xp) K *= dP[x];
    GemmPair& g = ep->h_pairs[p];
    std::memset(&g, 0, sizeof(g));
    g.off_a = ch.permA ? offA[bp] : P.offs[bp];
    g.off_b = ch.permB ? offB[bq] : Q.offs[bq];
    g.K = (int32_t)K;
    g.lda = (int32_t)((ep->ta && !ch.permA) ? std::max<int64_t>(K, 1) : std::max<int64_t>(M, 1));
    g.ldb = (int32_t)((ep->tb && !ch.permB) ? std::max<int64_t>(N, 1) : std::max<int64_t>(K, 1));
    alpha_of(p, &g.alpha_re, &g.alpha_im);
  }
  for (size_t gi = 0; gi < granges.size(); ++gi) {
    const size_t p = granges[gi].first;
    const int64_t br = pairs[p].blkR;
    const int64_t bp = swap ? pairs[p].blk2 : pairs[p].blk1;
    const int64_t bq = swap ? pairs[p].blk1 : pairs[p].blk2;
    int64_t M = 1, N = 1;
    for (int x : Mp) M *= P.d(bp)[x];
    for (int x : Nq) N *= Q.d(bq)[x];
    GemmGroup& g = ep->h_groups[gi];
    std::memset(&g, 0, sizeof(g));
    g.off_c = permC ? offC[br] : bR.offs[br];
    g.M = (int32_t)M;
    g.N = (int32_t)N;
    g.ldc = (int32_t)std::max<int64_t>(M, 1);
    g.p0 = (int32_t)granges[gi].first;
    g.p1 = (int32_t)granges[gi].second;
    double cost = 0;
    for (size_t q = granges[gi].first; q < granges[gi].second; ++q)
      cost += 2.0 * (double)M * (double)N * (double)ep->h_pairs[q].K;
    ep->group_cost[gi] = cost;
  }
  if ((e = ep->g_pairs.upload(ep->h_pairs.data(), ep->h_pairs.size() * sizeof(GemmPair), s)) !=
          hipSuccess ||
      (e = ep->g_groups.upload(ep->h_groups.data(), ep->h_groups.size() * sizeof(GemmGroup),
                               s)) != hipSuccess)
    return std::string("device table upload failed: ") + hipGetErrorString(e);
  std::string kerr = build_ksplit(ep, s);
  if (!kerr.empty()) return kerr;
  std::string serr = build_split_tables(ep, P, Q, Mp, Nq, ch.Xp, ch.Xq, pairs, swap, s);
  if (!serr.empty()) return serr;
  return "";
}

namespace {

// Mode list (positions in a block, in order, first fastest) -> merged
// (dims, strides) of a column-major block. Adjacent positions merge.
bool slice_modes(const std::vector<int>& pos, const int32_t* dims, int N, int32_t* mdim,
                 int32_t* mstr, int32_t* nm, int64_t* total) {
  int64_t st[kMaxModes];
  int64_t a = 1;
  for (int i = 0; i < N; ++i) {
    st[i] = a;
    a *= dims[i];
  }
  int n = 0;
  int64_t tot = 1;
  for (size_t t = 0; t < pos.size(); ++t) {
    const int p = pos[t];
    tot *= dims[p];
    if (n > 0 && t > 0 && pos[t - 1] + 1 == p) {
      mdim[n - 1] *= dims[p];
      continue;
    }
    if (n == kSliceModes) return false;
    mdim[n] = dims[p];
    mstr[n] = (int32_t)st[p];
    ++n;
  }
  *nm = n;
  *total = tot;
  return true;
}

}  // namespace

std::string build_split_tables(ExecPlan* ep, const BlockList& P, const BlockList& Q,
                               const std::vector<int>& Mp, const std::vector<int>& Nq,
                               const std::vector<int>& Xp, const std::vector<int>& Xq,
                               const std::vector<PairSpec>& pairs, bool swap, hipStream_t s) {
  ep->split_ok = false;
  auto side = [&](const BlockList& T, const std::vector<int>& rows, const std::vector<int>& ks,
                  std::vector<SliceBlock>* out, std::vector<int64_t>* prefix, int64_t* plane_bytes,
                  int64_t* nrows, int64_t* npmax) -> std::string {
    out->resize(T.nblk());
    prefix->assign(T.nblk() + 1, 0);
    int64_t pb = 0, er = 0, pm = 0;
    for (int64_t b = 0; b < T.nblk(); ++b) {
      SliceBlock& d = (*out)[b];
      std::memset(&d, 0, sizeof(d));
      int64_t R = 1, K = 1;
      for (int i = 0; i < kSliceModes; ++i) d.rdim[i] = d.kdim[i] = 1;
      if (!slice_modes(rows, T.d(b), T.N, d.rdim, d.rstr, &d.nrm, &R) ||
          !slice_modes(ks, T.d(b), T.N, d.kdim, d.kstr, &d.nkm, &K))
        return "too many modes for the split-precision operand reader";
      if (K > INT32_MAX / 2) return "contracted extent too large for the split-precision reader";
      d.off = T.offs[b];
      d.R = (int32_t)R;
      d.K = (int32_t)K;
      d.Kpad = (int32_t)((K + 31) / 32 * 32);
      d.kparts = (d.Kpad + kSliceKPart - 1) / kSliceKPart;
      d.plane_off = pb;
      d.exp_off = er;
      d.pmax_off = pm;
      pb += (int64_t)kSliceCount * R * d.Kpad;
      er += R;
      pm += (int64_t)d.kparts * R;
      (*prefix)[b + 1] = (*prefix)[b] + slice_block_wgs(d.R, d.kparts);
    }
    *plane_bytes = pb;
    *nrows = er;
    *npmax = pm;
    return "";
  };
  std::vector<SliceBlock> dA, dB;
  std::vector<int64_t> pA, pB;
  std::string err =
      side(P, Mp, Xp, &dA, &pA, &ep->sA_plane_bytes, &ep->sA_rows, &ep->sA_pmax);
  if (err.empty())
    err = side(Q, Nq, Xq, &dB, &pB, &ep->sB_plane_bytes, &ep->sB_rows, &ep->sB_pmax);
  if (!err.empty()) {
    ep->split_why = err;
    return "";  // the FP64 path still works
  }
  // One split pair per K chunk of at most kSplitKChunk (the INT32 level sums
  // of a chunk are exact; chunks of one output block are summed in FP64 like
  // pairs). Chunk [k0, k0 + len) of a block's planes starts k0 * R bytes in
  // (the planes are chunk-major). The groups get their own pair ranges.
  std::vector<SplitPair> sp;
  std::vector<int32_t> first(pairs.size() + 1, 0);
  for (size_t p = 0; p < pairs.size(); ++p) {
    const int64_t bp = swap ? pairs[p].blk2 : pairs[p].blk1;
    const int64_t bq = swap ? pairs[p].blk1 : pairs[p].blk2;
    SplitPair x;
    std::memset(&x, 0, sizeof(x));
    x.a_pstride = (int64_t)dA[bp].R * dA[bp].Kpad;
    x.b_pstride = (int64_t)dB[bq].R * dB[bq].Kpad;
    x.a_exp = (int32_t)dA[bp].exp_off;
    x.b_exp = (int32_t)dB[bq].exp_off;
    x.alpha = ep->h_pairs[p].alpha_re;
    x.alpha_im = ep->h_pairs[p].alpha_im;
    if (ep->h_pairs[p].alpha_im != 0.0) ep->cplx_pair_alpha = true;
    ep->split_kmax = std::max<int64_t>(ep->split_kmax, dA[bp].K);
    const int64_t kpad = dA[bp].K == 0 ? 0 : dA[bp].Kpad;
    first[p] = (int32_t)sp.size();
    int64_t k0 = 0;
    do {
      const int64_t len = std::min<int64_t>(kpad - k0, kSplitKChunk);
      x.a_planes = dA[bp].plane_off;
      x.b_planes = dB[bq].plane_off;
      x.a_koff = k0 * dA[bp].R;
      x.b_koff = k0 * dB[bq].R;
      x.kpad = (int32_t)len;
      sp.push_back(x);
      k0 += len;
    } while (k0 < kpad);
  }
  first[pairs.size()] = (int32_t)sp.size();
  std::vector<GemmGroup> sg(ep->h_groups);
  for (GemmGroup& G : sg) {
    G.p0 = first[G.p0];
    G.p1 = first[G.p1];
  }
  // one slicing grid over both operands: A's blocks, then B's (side 1)
  std::vector<SliceBlock> dAB(dA);
  std::vector<int64_t> pAB(pA);
  for (size_t b = 0; b < dB.size(); ++b) {
    SliceBlock d = dB[b];
    d.side = 1;
    dAB.push_back(d);
    pAB.push_back(pA.back() + pB[b + 1]);
  }
  // workgroup -> (block, tile within the block): one table read per
  // workgroup instead of a search over the blocks' workgroup prefix
  auto wgmap = [](const std::vector<int64_t>& prefix) {
    std::vector<int2> m((size_t)std::max<int64_t>(prefix.back(), 1), int2{0, 0});
    for (size_t b = 0; b + 1 < prefix.size(); ++b)
      for (int64_t w = prefix[b]; w < prefix[b + 1]; ++w)
        m[(size_t)w] = int2{(int)b, (int)(w - prefix[b])};
    return m;
  };
  const std::vector<int2> mAB = wgmap(pAB), mB = wgmap(pB);
  hipError_t e;
  ep->sA_blocks = (int64_t)dA.size();
  ep->sA_grid = pA.back();
  ep->sB_grid = pB.back();
  if (pAB.back() > INT32_MAX) return "slicing grid too large";
  if ((e = ep->s_desc.upload(dAB.data(), dAB.size() * sizeof(SliceBlock), s)) != hipSuccess ||
      (e = ep->s_wgmap.upload(mAB.data(), mAB.size() * sizeof(int2), s)) != hipSuccess ||
      (e = ep->s_wgmapB.upload(mB.data(), mB.size() * sizeof(int2), s)) != hipSuccess ||
      (e = ep->s_pairs.upload(sp.data(), sp.size() * sizeof(SplitPair), s)) != hipSuccess ||
      (e = ep->s_groups.upload(sg.data(), sg.size() * sizeof(GemmGroup), s)) != hipSuccess)
    return std::string("device table upload failed: ") + hipGetErrorString(e);
  ep->s_blocks = (int64_t)dAB.size();
  ep->s_grid = pAB.back();
  ep->split_ok = true;
  return "";
}

int default_gemm_cfg(const ExecPlan&, int) {
  // 64x64 tiles whatever the block size. Timed on MI355X against the other
  // shapes (128x128, 128x64, 64x128 real; 128x64 complex) on the two-site
  // network's GEMMs at chi = 1024 (one 5120 x 4096 x 1024 block), the
  // chi = 4096 U(1) blocks (about 410 on a side) and 4096^3 / 2048^3
  // products, 64x64 is the fastest or within 3 % of it, and the large tiles
  // are slower by up to 30 % (128x128 real) and 14 % (128x64 complex): four
  // waves per SIMD hide the fill latency that the larger register tiles
  // cannot.
  return 0;
}

int default_split_cfg(const ExecPlan& ep) {
  int64_t tiles = 0;
  for (const GemmGroup& G : ep.h_groups)
    if (G.M > 0 && G.N > 0) tiles += (int64_t)((G.M + 127) / 128) * ((G.N + 63) / 64);
  return tiles >= 2 * 256 ? 2 : 0;
}

int mixed_big_cfg(int cplx) { return cplx ? 2 : 1; }

// Key of a cached launch: the mixed choice's tile shapes depend on the
// element type (a plan is shared by real and complex tensors of one block
// structure), every other tile table does not.
int gemm_launch_key(int cfg, int cplx) {
  if (cfg == kCfgMixed) return kCfgMixed + (cplx ? 1 : 0);
  return gemm_cfg_ok(cfg, cplx) ? cfg : 0;
}

hipError_t ensure_gemm_launch(ExecPlan& ep, int cfg, int cplx, hipStream_t s, bool may_build) {
  if (cfg != kCfgMixed && !gemm_cfg_ok(cfg, cplx)) cfg = 0;
  const int key = gemm_launch_key(cfg, cplx);
  auto it = ep.launches.find(key);
  if (it != ep.launches.end()) return hipSuccess;
  if (!may_build) return hipErrorNotReady;
  std::unique_ptr<GemmLaunch> gl(new GemmLaunch);
  gl->cfg = cfg;
  hipError_t e;
  if (cfg == kCfgMixed) {
    const int big = mixed_big_cfg(cplx);
    std::vector<int2> t1 = xcd_tile_table(ep, gemm_cfg(big), 1);
    std::vector<int2> t2 = xcd_tile_table(ep, gemm_cfg(0), 2);
    gl->cfg = big;
    gl->grid = (int64_t)t1.size();
    gl->cfg2 = 0;
    gl->grid2 = (int64_t)t2.size();
    if ((e = gl->tiles.upload(t1.data(), t1.size() * sizeof(int2), s)) != hipSuccess) return e;
    if ((e = gl->tiles2.upload(t2.data(), t2.size() * sizeof(int2), s)) != hipSuccess) return e;
  } else {
    std::vector<int2> table = xcd_tile_table(ep, gemm_cfg(cfg));
    gl->grid = (int64_t)table.size();
    if ((e = gl->tiles.upload(table.data(), table.size() * sizeof(int2), s)) != hipSuccess)
      return e;
  }
  ep.launches[key] = std::move(gl);
  return hipSuccess;
}

// Workgroup w of a launch is dispatched to XCD w % kXcds. Cut the tiles of
// every group into chunks (a group's tiles in tm-fastest order; big groups
// are split so that every XCD gets several chunks), assign chunks to XCDs
// largest-work-first to the least loaded one, and interleave the per-XCD
// lists so each XCD runs its own chunks back to back: the tiles of a chunk
// share their A and B panels through that XCD's L2.
std::vector<int2> xcd_tile_table(const ExecPlan& ep, GemmTileCfg tc, int filter) {
  constexpr int kXcds = 8;
  struct Chunk {
    int32_t g, t0, t1;
    double work;
  };
  std::vector<Chunk> chunks;
  int64_t total = 0;
  std::vector<int32_t> gt(ep.h_groups.size());
  for (size_t g = 0; g < ep.h_groups.size(); ++g) {
    const GemmGroup& G = ep.h_groups[g];
    const bool big = G.M >= tc.bm && G.N >= tc.bn;
    const bool keep = filter == 0 || (filter == 1 && big) || (filter == 2 && !big);
    gt[g] = (!keep || G.M <= 0 || G.N <= 0)
                ? 0
                : ((G.M + tc.bm - 1) / tc.bm) * ((G.N + tc.bn - 1) / tc.bn);
    total += gt[g];
  }
  const int64_t unit = std::max<int64_t>(1, (total + 8 * kXcds - 1) / (8 * kXcds));
  for (size_t g = 0; g < ep.h_groups.size(); ++g) {
    double k = 0;
    for (int p = ep.h_groups[g].p0; p < ep.h_groups[g].p1; ++p) k += std::max(ep.h_pairs[p].K, 1);
    for (int64_t t0 = 0; t0 < gt[g]; t0 += unit) {
      const int64_t t1 = std::min<int64_t>(gt[g], t0 + unit);
      chunks.push_back(Chunk{(int32_t)g, (int32_t)t0, (int32_t)t1, k * (double)(t1 - t0)});
    }
  }
  std::stable_sort(chunks.begin(), chunks.end(),
                   [](const Chunk& a, const Chunk& b) { return a.work > b.work; });
  std::vector<std::vector<int2>> bins(kXcds);
  std::vector<double> load(kXcds, 0.0);
  for (const Chunk& c : chunks) {
    int best = 0;
    for (int x = 1; x < kXcds; ++x)
      if (load[x] < load[best]) best = x;
    load[best] += c.work;
    for (int32_t t = c.t0; t < c.t1; ++t) bins[best].push_back(int2{c.g, t});
  }
  size_t len = 0;
  for (auto& b : bins) len = std::max(len, b.size());
  std::vector<int2> table(len * kXcds, int2{-1, -1});
  for (int x = 0; x < kXcds; ++x)
    for (size_t i = 0; i < bins[x].size(); ++i) table[i * kXcds + x] = bins[x][i];
  return table;
}

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

// Complex split products whose scale is not real (a complex alpha or
// complex per-pair factors) are formed in separate buffers and combined.
static bool split_cplx_scale(const ExecPlan& ep, int cplx, bool alpha_cplx) {
  return cplx && (alpha_cplx || ep.cplx_pair_alpha);
}

// Split precision hands a plan to the FP64 split-K form when one of its
// output blocks sums over at least this many contracted values: the split
// GEMM has no K split, and such a product is a reduction, not GEMM-shaped
// work.
constexpr int64_t kKsSplitPrecK = 1 << 17;
bool ksplit_takes_split_prec(const ExecPlan& ep) {
  return ep.ks_grid > 0 && ep.ks_maxk >= kKsSplitPrecK;
}

size_t exec_workspace_bytes(const ExecPlan& ep, int cplx, int prec, bool alpha_cplx) {
  if (ep.strategy != ExecPlan::GEMM) return 0;
  const size_t es = cplx ? 16 : 8;
  size_t b = 0;
  if (prec != 0 && ksplit_takes_split_prec(ep)) prec = 0;
  if (prec == 0 && ep.ks_grid > 0) b += align256((size_t)ep.ks_elems * es);
  if (prec != 0 && ep.split_ok) {
    const size_t ns = cplx ? 3 : 1;
    b += align256((size_t)ep.sA_plane_bytes * ns) + align256((size_t)ep.sB_plane_bytes * ns);
    b += align256((size_t)ep.sA_rows * 4) + align256((size_t)ep.sB_rows * 4);
    b += align256((size_t)ep.sA_pmax * 4) + align256((size_t)ep.sB_pmax * 4);
    if (split_cplx_scale(ep, cplx, alpha_cplx)) {
      // sum of the real-factor parts, then (complex pair factors) the
      // imaginary-factor parts, each in C's (pre-permutation) layout
      b += align256((size_t)ep.nnzC * es) * (ep.cplx_pair_alpha ? 2 : 1);
      if (ep.permC) return b + 1024;  // the combined sum is permuted into R from there
    }
  } else {
    if (ep.permA) b += align256((size_t)ep.nnzA * es);
    if (ep.permB) b += align256((size_t)ep.nnzB * es);
  }
  if (ep.permC) b += align256((size_t)ep.nnzC * es);
  return b + 1024;
}

namespace {

// Slice sets of the two K halves of a complex split product (see
// launch_split_gemm): Re C = Re A Re B - Im A Im B, Im C = Re A Im B + Im A Re B,
// with Im A -> -Im A under conj(A) and Im B -> -Im B under conj(B).
int split_sets(int comp, bool conja, bool conjb) {
  const int ima = conja ? 2 : 1, imb = conjb ? 2 : 1, nimb = conjb ? 1 : 2;
  int a0, b0, a1, b1;
  if (comp == 0) {
    a0 = 0, b0 = 0, a1 = ima, b1 = nimb;
  } else {
    a0 = 0, b0 = imb, a1 = ima, b1 = 0;
  }
  return a0 | (b0 << 2) | (a1 << 4) | (b1 << 6);
}

// Brackets the launches made in its lifetime with timer events (no-op when
// the timer is null).
struct TimeScope {
  KernelTimer* kt;
  hipStream_t s;
  TimeScope(KernelTimer* k, const char* name, hipStream_t st) : kt(k), s(st) {
    if (kt) kt->begin(name, s);
  }
  ~TimeScope() {
    if (kt) kt->end(s);
  }
};

hipError_t exec_split(ExecPlan& ep, int cplx, int prec, int cfg, hipStream_t s, void* ws,
                      const void* T1, const void* T2, void* R, unsigned flags, double a_re,
                      double a_im, double b_re, double b_im, std::string* err,
                      KernelTimer* kt, bool may_build) {
  if (!ep.split_ok) {
    *err = "split precision is unavailable for this contraction: " + ep.split_why;
    return hipErrorNotSupported;
  }
  if (split_slices(prec) == 0) {
    *err = "this split-precision form is not available";
    return hipErrorNotSupported;
  }
  const int nsets = cplx ? 3 : 1;
  const void* A = ep.swap ? T2 : T1;
  const void* B = ep.swap ? T1 : T2;
  const bool conja = (flags & (ep.swap ? 2u : 1u)) != 0;
  const bool conjb = (flags & (ep.swap ? 1u : 2u)) != 0;
  char* w = (char*)ws;
  void* planesA = w;
  w += align256((size_t)ep.sA_plane_bytes * nsets);
  void* planesB = w;
  w += align256((size_t)ep.sB_plane_bytes * nsets);
  void* expA = w;
  w += align256((size_t)ep.sA_rows * 4);
  void* expB = w;
  w += align256((size_t)ep.sB_rows * 4);
  void* pmaxA = w;
  w += align256((size_t)ep.sA_pmax * 4);
  void* pmaxB = w;
  w += align256((size_t)ep.sB_pmax * 4);
  hipError_t e;
  // fixed operands: side 0 is A (T1 unless swapped), side 1 is B. A kept
  // operand's planes and exponents live in the plan's own buffer; a reused
  // one (checked by the caller: kept from the same source with the same set
  // count) is not sliced again.
  const unsigned fk1 = flags & kFlagKeep1, fk2 = flags & kFlagKeep2;
  const unsigned fr1 = flags & kFlagReuse1, fr2 = flags & kFlagReuse2;
  const bool keep[2] = {(ep.swap ? fk2 : fk1) != 0, (ep.swap ? fk1 : fk2) != 0};
  const bool reuse[2] = {!keep[0] && (ep.swap ? fr2 : fr1) != 0,
                         !keep[1] && (ep.swap ? fr1 : fr2) != 0};
  void** planes_of[2] = {&planesA, &planesB};
  void** exps_of[2] = {&expA, &expB};
  const int64_t pbytes[2] = {ep.sA_plane_bytes, ep.sB_plane_bytes};
  const int64_t rows[2] = {ep.sA_rows, ep.sB_rows};
  for (int side = 0; side < 2; ++side) {
    if (!keep[side] && !reuse[side]) continue;
    const size_t pb = align256((size_t)pbytes[side] * nsets);
    if (keep[side]) {
      if ((e = ep.kept[side].reserve(pb + align256((size_t)rows[side] * 4))) != hipSuccess) {
        *err = std::string("kept slice storage: ") + hipGetErrorString(e);
        return e;
      }
      ep.kept_sets[side] = 0;  // valid again once sliced below
    }
    *planes_of[side] = ep.kept[side].p;
    *exps_of[side] = (char*)ep.kept[side].p + pb;
  }
  if (!reuse[0] || !reuse[1]) {
    SliceOperands ops;
    ops.src[0] = (const double*)A;
    ops.src[1] = (const double*)B;
    ops.planes[0] = (int8_t*)planesA;
    ops.planes[1] = (int8_t*)planesB;
    ops.exps[0] = (int32_t*)expA;
    ops.exps[1] = (int32_t*)expB;
    ops.pmax[0] = (int32_t*)pmaxA;
    ops.pmax[1] = (int32_t*)pmaxB;
    const SliceBlock* desc = (const SliceBlock*)ep.s_desc.p;
    TimeScope ts(kt, "slice_max_kernel+slice_kernel", s);
    if (!reuse[0] && !reuse[1])
      e = launch_slice(s, prec, cplx, nsets, desc, (const int2*)ep.s_wgmap.p, ep.s_grid, ops);
    else if (!reuse[0])
      e = launch_slice(s, prec, cplx, nsets, desc, (const int2*)ep.s_wgmap.p, ep.sA_grid, ops);
    else
      e = launch_slice(s, prec, cplx, nsets, desc + ep.sA_blocks, (const int2*)ep.s_wgmapB.p,
                       ep.sB_grid, ops);
    if (e != hipSuccess) {
      *err = std::string("slice launch: ") + hipGetErrorString(e);
      return e;
    }
  }
  for (int side = 0; side < 2; ++side)
    if (keep[side]) {
      ep.kept_src[side] = side == 0 ? A : B;
      ep.kept_sets[side] = nsets;
    }
  // cfg 2: 128 x 64 split tiles (three-stage); cfg 3 (real data): 64 x 64
  // with A in registers; otherwise 64 x 64
  const int wmb = cfg == 2 ? 2 : (cfg == 3 && !cplx ? 3 : 1);
  const int tcfg = wmb == 2 ? 2 : 0;  // the FP64 tile table of the same shape
  if ((e = ensure_gemm_launch(ep, tcfg, 0, s, may_build)) != hipSuccess) {
    *err = e == hipErrorNotReady ? std::string(kTablesMissing)
                                 : std::string("gemm table upload: ") + hipGetErrorString(e);
    return e;
  }
  GemmLaunch& gl = *ep.launches[tcfg];
  if (split_cplx_scale(ep, cplx, a_im != 0.0)) {
    // R = alpha * sum_p (a_p + i b_p) P_p + beta * R: the split kernel scales
    // each pair by a real factor, so form S1 = sum_p a_p P_p and S2 =
    // sum_p b_p P_p (unit global scale), then R = alpha (S1 + i S2) + beta R
    // (through the output permutation when there is one)
    const size_t cbytes = align256((size_t)ep.nnzC * 16);
    void* S1 = w;
    void* S2 = ep.cplx_pair_alpha ? (void*)(w + cbytes) : nullptr;
    for (int part = 0; part < (S2 ? 2 : 1) && e == hipSuccess; ++part)
      for (int comp = 0; comp < 2 && e == hipSuccess; ++comp) {
        TimeScope ts(kt, "split_gemm_kernel", s);
        e = launch_split_gemm(s, prec, cplx, wmb, comp, split_sets(comp, conja, conjb),
                              (const SplitPair*)ep.s_pairs.p, (const GemmGroup*)ep.s_groups.p,
                              (const int2*)gl.tiles.p, gl.grid, planesA, planesB, expA, expB,
                              part ? S2 : S1, 1.0, 0.0, part);
      }
    if (e != hipSuccess) {
      *err = std::string("split gemm launch: ") + hipGetErrorString(e);
      return e;
    }
    if (ep.permC) {
      if (S2) {
        TimeScope ts(kt, "axpby_join_kernel", s);
        e = launch_axpby_join(s, ep.nnzC, 1.0, 0.0, S1, S2, 0.0, 0.0, S1);
        if (e != hipSuccess) {
          *err = std::string("axpby launch: ") + hipGetErrorString(e);
          return e;
        }
      }
      TimeScope ts(kt, "permute_kernel", s);
      e = launch_permute(s, cplx, ep.pC.nm, ep.pC.perm, (const PermBlock*)ep.pC.blocks.p,
                         (const int64_t*)ep.pC.prefix.p, ep.pC.nblk, ep.pC.tiles, ep.pC.kind, S1,
                         R, a_re, a_im, b_re, b_im);
      if (e != hipSuccess) *err = std::string("permute launch: ") + hipGetErrorString(e);
      return e;
    }
    TimeScope ts(kt, "axpby_join_kernel", s);
    e = launch_axpby_join(s, ep.nnzC, a_re, a_im, S1, S2, b_re, b_im, R);
    if (e != hipSuccess) *err = std::string("axpby launch: ") + hipGetErrorString(e);
    return e;
  }
  void* C = ep.permC ? (void*)w : R;
  double beta_k = ep.permC ? 0.0 : b_re;
  if (cplx && !ep.permC && b_im != 0.0) {
    // complex beta mixes the components: scale R first, then accumulate
    {
      TimeScope ts(kt, "scale_kernel", s);
      e = launch_scale(s, 1, R, ep.nnzC, b_re, b_im);
    }
    if (e != hipSuccess) {
      *err = std::string("scale launch: ") + hipGetErrorString(e);
      return e;
    }
    beta_k = 1.0;
  }
  for (int comp = 0; comp < (cplx ? 2 : 1) && e == hipSuccess; ++comp) {
    TimeScope ts(kt, "split_gemm_kernel", s);
    e = launch_split_gemm(s, prec, cplx, wmb, comp, split_sets(comp, conja, conjb),
                          (const SplitPair*)ep.s_pairs.p, (const GemmGroup*)ep.s_groups.p,
                          (const int2*)gl.tiles.p, gl.grid, planesA, planesB, expA, expB, C, a_re,
                          beta_k);
  }
  if (e != hipSuccess) {
    *err = std::string("split gemm launch: ") + hipGetErrorString(e);
    return e;
  }
  if (ep.permC) {
    TimeScope ts(kt, "permute_kernel", s);
    e = launch_permute(s, cplx, ep.pC.nm, ep.pC.perm, (const PermBlock*)ep.pC.blocks.p,
                       (const int64_t*)ep.pC.prefix.p, ep.pC.nblk, ep.pC.tiles, ep.pC.kind, C, R,
                       1.0, 0.0, b_re, b_im);
    if (e != hipSuccess) *err = std::string("permute launch: ") + hipGetErrorString(e);
  }
  return e;
}

}  // namespace

hipError_t exec_run(ExecPlan& ep, int cplx, int prec, int cfg, hipStream_t s, void* ws,
                    const void* T1, const void* T2, void* R, unsigned flags, double a_re,
                    double a_im, double b_re, double b_im, std::string* err, KernelTimer* kt,
                    bool may_build) {
  hipError_t e;
  if (ep.strategy == ExecPlan::DIRECT) {
    {
      TimeScope ts(kt, "direct_kernel", s);
      e = launch_direct(s, cplx, ep.st.NR, (const DirectPair*)ep.d_pairs.p,
                        (const DirectGroup*)ep.d_groups.p, (const int32_t*)ep.d_koffs.p,
                        (const int32_t*)ep.d_wgg.p, ep.d_tiles, T1, T2, R, flags & kFlagConj, a_re, a_im,
                        b_re, b_im);
    }
    if (e != hipSuccess) *err = std::string("direct kernel launch: ") + hipGetErrorString(e);
    return e;
  }
  if (ep.strategy == ExecPlan::FIBER) {
    const void* A = ep.f_a_is_2 ? T2 : T1;
    const void* B = ep.f_a_is_2 ? T1 : T2;
    const unsigned fl =
        ep.f_a_is_2 ? (((flags & 1u) << 1) | ((flags & 2u) >> 1)) : (flags & kFlagConj);
    if (ep.f_mv) {
      TimeScope ts(kt, "fiber_mv_kernel", s);
      e = launch_fiber_mv(s, cplx, ep.f_nf, ep.f_smax, (const FiberMVGroup*)ep.fm_groups.p,
                          (const FiberSlot*)ep.fm_slots.p, (const FiberOutElem*)ep.fm_outs.p,
                          (const FiberCoef*)ep.fm_coefs.p, (const int32_t*)ep.f_wgg.p, ep.f_wgs,
                          A, B, R, fl, a_re, a_im, b_re, b_im);
    } else {
      TimeScope ts(kt, "fiber_kernel", s);
      e = launch_fiber(s, cplx, ep.f_nf, (const FiberGroup*)ep.f_groups.p,
                       (const FiberOut*)ep.f_outs.p, (const FiberPair*)ep.f_pairs.p,
                       (const int32_t*)ep.f_koffs.p, (const int32_t*)ep.f_gofs_c.p,
                       (const int32_t*)ep.f_gofs_b.p, (const int32_t*)ep.f_wgg.p, ep.f_wgs, A, B,
                       R, fl, a_re, a_im, b_re, b_im);
    }
    if (e != hipSuccess) *err = std::string("fiber kernel launch: ") + hipGetErrorString(e);
    return e;
  }
  if (prec != 0 && !ksplit_takes_split_prec(ep))
    return exec_split(ep, cplx, prec, cfg, s, ws, T1, T2, R, flags, a_re, a_im, b_re, b_im, err,
                      kt, may_build);
  const size_t es = cplx ? 16 : 8;
  char* w = (char*)ws;
  const void* A = ep.swap ? T2 : T1;
  const void* B = ep.swap ? T1 : T2;
  unsigned gflags = ep.swap ? (((flags & 1u) << 1) | ((flags & 2u) >> 1)) : (flags & kFlagConj);
  if (ep.permA) {
    {
      TimeScope ts(kt, "permute_kernel", s);
      e = launch_permute(s, cplx, ep.pA.nm, ep.pA.perm, (const PermBlock*)ep.pA.blocks.p,
                         (const int64_t*)ep.pA.prefix.p, ep.pA.nblk, ep.pA.tiles, ep.pA.kind, A,
                         w, 1.0, 0.0, 0.0, 0.0);
    }
    if (e != hipSuccess) {
      *err = std::string("permute launch: ") + hipGetErrorString(e);
      return e;
    }
    A = w;
    w += (size_t)ep.nnzA * es;
  }
  if (ep.permB) {
    {
      TimeScope ts(kt, "permute_kernel", s);
      e = launch_permute(s, cplx, ep.pB.nm, ep.pB.perm, (const PermBlock*)ep.pB.blocks.p,
                         (const int64_t*)ep.pB.prefix.p, ep.pB.nblk, ep.pB.tiles, ep.pB.kind, B,
                         w, 1.0, 0.0, 0.0, 0.0);
    }
    if (e != hipSuccess) {
      *err = std::string("permute launch: ") + hipGetErrorString(e);
      return e;
    }
    B = w;
    w += (size_t)ep.nnzB * es;
  }
  void* C = R;
  double gbr = b_re, gbi = b_im;
  if (ep.permC) {
    C = w;
    w += (size_t)ep.nnzC * es;
    gbr = 0.0;
    gbi = 0.0;
  }
  if (ep.ks_grid > 0) {
    // split-K: partial blocks into the workspace, then their sum into C
    if (ep.ks_dot) {
      TimeScope ts(kt, "ksplit_dot_kernel", s);
      e = launch_ksplit_dot(s, cplx, ep.ks_dot1, ep.ta, ep.tb, (const GemmPair*)ep.ks_pairs.p,
                            (const GemmGroup*)ep.ks_groups.p, ep.ks_ngroups, A, B, w, gflags,
                            a_re, a_im);
    } else if (ep.ks_skinny) {
      TimeScope ts(kt, "ksplit_skinny_kernel", s);
      e = launch_ksplit_skinny(s, cplx, ep.ks_kc, ep.ta, ep.tb, (const GemmPair*)ep.ks_pairs.p,
                               (const GemmGroup*)ep.ks_groups.p, (const int2*)ep.ks_rows.p,
                               ep.ks_nrows, A, B, w, gflags, a_re, a_im);
    } else {
      TimeScope ts(kt, "gemm_f64_kernel", s);
      e = launch_gemm_f64(s, 0, cplx, ep.ta, ep.tb, (const GemmPair*)ep.ks_pairs.p,
                          (const GemmGroup*)ep.ks_groups.p, (const int2*)ep.ks_tiles.p,
                          ep.ks_grid, A, B, w, gflags, a_re, a_im, 0.0, 0.0);
    }
    if (e == hipSuccess) {
      TimeScope ts(kt, "ksplit_reduce_kernel", s);
      e = launch_ksplit_reduce(s, cplx, (const KsReduce*)ep.ks_red.p, ep.ks_nred, ep.ks_total,
                               ep.ks_maxs, w, C, gbr, gbi);
    }
    if (e != hipSuccess) {
      *err = std::string("split-K gemm launch: ") + hipGetErrorString(e);
      return e;
    }
  } else {
  if ((e = ensure_gemm_launch(ep, cfg, cplx, s, may_build)) != hipSuccess) {
    *err = e == hipErrorNotReady ? std::string(kTablesMissing)
                                 : std::string("gemm table upload: ") + hipGetErrorString(e);
    return e;
  }
  GemmLaunch& gl = *ep.launches[gemm_launch_key(cfg, cplx)];
  {
    TimeScope ts(kt, "gemm_f64_kernel", s);
    e = launch_gemm_f64(s, gl.cfg, cplx, ep.ta, ep.tb, (const GemmPair*)ep.g_pairs.p,
                        (const GemmGroup*)ep.g_groups.p, (const int2*)gl.tiles.p, gl.grid, A, B,
                        C, gflags, a_re, a_im, gbr, gbi);
  }
  if (e == hipSuccess && gl.cfg2 >= 0) {
    TimeScope ts(kt, "gemm_f64_kernel", s);
    e = launch_gemm_f64(s, gl.cfg2, cplx, ep.ta, ep.tb, (const GemmPair*)ep.g_pairs.p,
                        (const GemmGroup*)ep.g_groups.p, (const int2*)gl.tiles2.p, gl.grid2, A, B,
                        C, gflags, a_re, a_im, gbr, gbi);
  }
  if (e != hipSuccess) {
    char what[160];
    std::snprintf(what, sizeof what,
                  "gemm launch (tile cfg %d / %d, workgroups %lld / %lld, %s, ta %d, tb %d): ",
                  gl.cfg, gl.cfg2, (long long)gl.grid, (long long)gl.grid2,
                  cplx ? "complex" : "real", ep.ta, ep.tb);
    *err = std::string(what) + hipGetErrorString(e);
    return e;
  }
  }  // split-K
  if (ep.permC) {
    TimeScope ts(kt, "permute_kernel", s);
    e = launch_permute(s, cplx, ep.pC.nm, ep.pC.perm, (const PermBlock*)ep.pC.blocks.p,
                       (const int64_t*)ep.pC.prefix.p, ep.pC.nblk, ep.pC.tiles, ep.pC.kind, C, R,
                       1.0, 0.0, b_re, b_im);
    if (e != hipSuccess) *err = std::string("permute launch: ") + hipGetErrorString(e);
  }
  return e;
}

// ---------------------------------------------------------------- timer
hipEvent_t KernelTimer::take() {
  if (!pool.empty()) {
    hipEvent_t ev = pool.back();
    pool.pop_back();
    return ev;
  }
  hipEvent_t ev = nullptr;
  if (hipEventCreate(&ev) != hipSuccess) return nullptr;
  return ev;
}

LaunchEvents& launch_events() {
  thread_local LaunchEvents ev;
  return ev;
}

// The section's kernels record the two events themselves (launch_kernel):
// start of the first dispatch, end of the last, the interval rocprofv3
// reports for a single kernel.
void KernelTimer::begin(const char* name, hipStream_t) {
  open_name = name;
  open_a = take();
  open_b = take();
  LaunchEvents& ev = launch_events();
  ev = LaunchEvents{};
  if (open_a && open_b) {
    ev.start = open_a;
    ev.stop = open_b;
  }
}

void KernelTimer::end(hipStream_t) {
  LaunchEvents& ev = launch_events();
  const bool used = ev.launched;
  ev = LaunchEvents{};
  if (!open_a || !open_b || !used) {  // nothing launched: nothing to time
    if (open_a) pool.push_back(open_a);
    if (open_b) pool.push_back(open_b);
  } else {
    pending.push_back(Pending{open_name, open_a, open_b});
  }
  open_a = open_b = nullptr;
}

hipError_t KernelTimer::fold() {
  hipError_t first = hipSuccess;
  for (Pending& p : pending) {
    hipError_t e = hipEventSynchronize(p.b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.a, p.b);
    if (e == hipSuccess) {
      auto& t = totals[p.name];
      t.first += 1;
      t.second += ms;
    } else if (first == hipSuccess) {
      first = e;
    }
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  pending.clear();
  return first;
}

void KernelTimer::clear() {
  (void)fold();
  totals.clear();
}

KernelTimer::~KernelTimer() {
  for (Pending& p : pending) {
    (void)hipEventSynchronize(p.b);
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  for (hipEvent_t ev : pool) (void)hipEventDestroy(ev);
}

// Fused pair of matrix-form fiber steps (ChainGroup). Returns "" when
// built, "-" when the pair has no fused form, or an error message.
std::string build_fiber_chain(const ExecPlan& e1, const ExecPlan& e2, hipStream_t s,
                              ChainPlan* cp) {
  if (e1.strategy != ExecPlan::FIBER || e2.strategy != ExecPlan::FIBER || !e1.f_mv ||
      !e2.f_mv || e1.f_a_is_2 || e2.f_a_is_2)
    return "-";
  const int nf1 = (int)e1.f_fm.size(), nf2 = (int)e2.f_fm.size();
  // Y modes step 1 iterates over as fibers (its output modes) and those
  // step 2 iterates over (its A modes); Y is step 1's output and step 2's A
  std::vector<int> p1(kMaxModes * 2, -1), p2(kMaxModes * 2, -1);
  for (int i = 0; i < nf1; ++i) p1[(size_t)e1.f_fm[(size_t)i].second] = i;
  for (int i = 0; i < nf2; ++i) p2[(size_t)e2.f_fm[(size_t)i].first] = i;
  std::vector<int> i1, i2;  // fiber positions of the shared modes in each step
  for (int i = 0; i < nf1; ++i) {
    const int y = e1.f_fm[(size_t)i].second;
    if (p2[(size_t)y] >= 0) {
      i1.push_back(i);
      i2.push_back(p2[(size_t)y]);
    }
  }
  const int ns = (int)i1.size();
  if (ns > 4) return "-";
  // the modes only one step iterates over must have unit extents
  for (const FiberMVGroup& g : e1.hfm_groups)
    for (int i = 0; i < nf1; ++i)
      if (p2[(size_t)e1.f_fm[(size_t)i].second] < 0 && g.fdim[i] != 1) return "-";
  for (const FiberMVGroup& g : e2.hfm_groups)
    for (int i = 0; i < nf2; ++i)
      if (p1[(size_t)e2.f_fm[(size_t)i].first] < 0 && g.fdim[i] != 1) return "-";

  // Y element (offset at fiber 0) -> (step-1 group, its output)
  const size_t n1 = e1.hfm_groups.size(), n2 = e2.hfm_groups.size();
  std::unordered_map<int64_t, std::pair<int32_t, int32_t>> prod;
  for (size_t g = 0; g < n1; ++g) {
    const FiberMVGroup& G = e1.hfm_groups[g];
    for (int o = 0; o < G.nout; ++o) {
      const FiberOutElem& oe = e1.hfm_outs[(size_t)(G.o0 + o)];
      if (!prod.emplace(oe.off_c + oe.go, std::make_pair((int32_t)g, (int32_t)o)).second)
        return "-";  // written twice
    }
  }
  std::vector<size_t> par(n1 + n2);
  std::iota(par.begin(), par.end(), 0);
  auto find = [&](size_t x) {
    while (par[x] != x) x = par[x] = par[par[x]];
    return x;
  };
  for (size_t h = 0; h < n2; ++h) {
    const FiberMVGroup& H = e2.hfm_groups[h];
    for (int t = 0; t < H.nslot; ++t) {
      const FiberSlot& sl = e2.hfm_slots[(size_t)(H.s0 + t)];
      auto it = prod.find(sl.off_a + sl.ka);
      if (it == prod.end()) return "-";  // a Y value step 1 does not write
      const FiberMVGroup& G = e1.hfm_groups[(size_t)it->second.first];
      const FiberOutElem& oe = e1.hfm_outs[(size_t)(G.o0 + it->second.second)];
      for (int j = 0; j < ns; ++j)  // the same Y element at every super-fiber
        if (oe.cstr[i1[(size_t)j]] != sl.astr[i2[(size_t)j]] ||
            G.fdim[i1[(size_t)j]] != H.fdim[i2[(size_t)j]])
          return "-";
      par[find((size_t)it->second.first)] = find(n1 + h);
    }
  }
  std::map<size_t, std::pair<std::vector<size_t>, std::vector<size_t>>> comps;
  for (size_t g = 0; g < n1; ++g) comps[find(g)].first.push_back(g);
  for (size_t h = 0; h < n2; ++h) comps[find(n1 + h)].second.push_back(h);

  std::vector<ChainGroup> groups;
  std::vector<FiberSlot> slots;
  std::vector<FiberOutElem> zouts;
  std::vector<uint32_t> ymask;
  std::vector<FiberCoef> coefs1, coefs2;
  std::vector<int32_t> wgg;
  int64_t wgs = 0;
  int xmax = 1, ymax = 1;
  for (auto& kv : comps) {
    const std::vector<size_t>& G1 = kv.second.first;
    const std::vector<size_t>& G2 = kv.second.second;
    if (G2.empty()) continue;  // Y values nothing downstream reads
    if (G1.empty()) return "-";
    ChainGroup C;
    std::memset(&C, 0, sizeof(C));
    const FiberMVGroup& g0 = e1.hfm_groups[G1[0]];
    C.nfib = g0.nfib;
    for (int i = 0; i < kMaxModes; ++i) C.fdim[i] = 1;
    for (int j = 0; j < ns; ++j) C.fdim[j] = g0.fdim[i1[(size_t)j]];
    for (size_t g : G1)
      for (int j = 0; j < ns; ++j)
        if (e1.hfm_groups[g].fdim[i1[(size_t)j]] != C.fdim[j]) return "-";
    // x: every step-1 slot, groups in order, strides over the shared modes
    C.x0 = (int32_t)slots.size();
    std::map<size_t, int32_t> xbase;
    for (size_t g : G1) {
      const FiberMVGroup& G = e1.hfm_groups[g];
      xbase[g] = (int32_t)(slots.size() - (size_t)C.x0);
      for (int u = 0; u < G.nslot; ++u) {
        const FiberSlot& s0 = e1.hfm_slots[(size_t)(G.s0 + u)];
        FiberSlot sl;
        std::memset(&sl, 0, sizeof(sl));
        sl.off_a = s0.off_a;
        sl.ka = s0.ka;
        for (int j = 0; j < ns; ++j) sl.astr[j] = s0.astr[i1[(size_t)j]];
        slots.push_back(sl);
      }
    }
    C.nx = (int32_t)(slots.size() - (size_t)C.x0);
    if (C.nx == 0) return "-";
    // y: every step-2 slot, groups in order, and the step-1 output writing it
    C.y0 = (int32_t)ymask.size();
    std::map<size_t, int32_t> ybase;
    std::map<std::pair<int32_t, int32_t>, int32_t> yof;
    for (size_t h : G2) {
      const FiberMVGroup& H = e2.hfm_groups[h];
      ybase[h] = (int32_t)(ymask.size() - (size_t)C.y0);
      for (int t = 0; t < H.nslot; ++t) {
        const FiberSlot& sl = e2.hfm_slots[(size_t)(H.s0 + t)];
        if (!yof.emplace(prod.at(sl.off_a + sl.ka), (int32_t)(ymask.size() - (size_t)C.y0))
                 .second)
          return "-";  // one Y value in two step-2 slots
        ymask.push_back(0u);
      }
    }
    C.ny = (int32_t)(ymask.size() - (size_t)C.y0);
    if (C.nx > kChainX || C.ny > kChainY) return "-";
    // W1: the step-1 coefficients of the Y values read downstream
    C.w10 = (int32_t)coefs1.size();
    for (size_t g : G1) {
      const FiberMVGroup& G = e1.hfm_groups[g];
      for (int i = 0; i < G.nw; ++i) {
        FiberCoef cf = e1.hfm_coefs[(size_t)(G.w0 + i)];
        auto it = yof.find({(int32_t)g, cf.out});
        if (it == yof.end()) continue;
        cf.slot = xbase[g] + cf.slot;
        cf.out = it->second;
        ymask[(size_t)(C.y0 + cf.out)] |= 1u << cf.slot;
        coefs1.push_back(cf);
      }
    }
    C.nw1 = (int32_t)(coefs1.size() - (size_t)C.w10);
    // z: every step-2 output (masks over y, strides over the shared modes); W2
    C.z0 = (int32_t)zouts.size();
    C.w20 = (int32_t)coefs2.size();
    for (size_t h : G2) {
      const FiberMVGroup& H = e2.hfm_groups[h];
      const int32_t zb = (int32_t)(zouts.size() - (size_t)C.z0);
      for (int o = 0; o < H.nout; ++o) {
        const FiberOutElem& o0 = e2.hfm_outs[(size_t)(H.o0 + o)];
        FiberOutElem oe;
        std::memset(&oe, 0, sizeof(oe));
        oe.off_c = o0.off_c;
        oe.go = o0.go;
        for (int j = 0; j < ns; ++j) oe.cstr[j] = o0.cstr[i2[(size_t)j]];
        for (int t = 0; t < H.nslot; ++t)
          if ((o0.mask >> t) & 1u) oe.mask |= 1u << (ybase[h] + t);
        zouts.push_back(oe);
      }
      for (int i = 0; i < H.nw; ++i) {
        FiberCoef cf = e2.hfm_coefs[(size_t)(H.w0 + i)];
        cf.slot = ybase[h] + cf.slot;
        cf.out = zb + cf.out;
        coefs2.push_back(cf);
      }
    }
    C.nz = (int32_t)(zouts.size() - (size_t)C.z0);
    C.nw2 = (int32_t)(coefs2.size() - (size_t)C.w20);
    if (C.nz == 0 || C.nz > kChainZ || C.ny == 0) return "-";
    xmax = std::max(xmax, (int)C.nx);
    ymax = std::max(ymax, (int)C.ny);
    C.wg0 = wgs;
    const int64_t n = (C.nfib + kFibersPerWGChain - 1) / kFibersPerWGChain;
    for (int64_t w = 0; w < n; ++w) wgg.push_back((int32_t)groups.size());
    wgs += n;
    groups.push_back(C);
  }
  if (groups.empty()) return "-";
  if (coefs1.empty()) coefs1.push_back(FiberCoef{});
  if (coefs2.empty()) coefs2.push_back(FiberCoef{});
  cp->nf = std::max(ns, 1);
  cp->xm = xmax <= 8 ? 8 : kChainX;
  cp->ym = ymax <= 8 ? 8 : kChainY;
  cp->wgs = wgs;
  hipError_t e;
  if ((e = cp->groups.upload(groups.data(), groups.size() * sizeof(ChainGroup), s)) != hipSuccess ||
      (e = cp->slots.upload(slots.data(), slots.size() * sizeof(FiberSlot), s)) != hipSuccess ||
      (e = cp->zouts.upload(zouts.data(), zouts.size() * sizeof(FiberOutElem), s)) !=
          hipSuccess ||
      (e = cp->ymask.upload(ymask.data(), ymask.size() * sizeof(uint32_t), s)) != hipSuccess ||
      (e = cp->coefs1.upload(coefs1.data(), coefs1.size() * sizeof(FiberCoef), s)) !=
          hipSuccess ||
      (e = cp->coefs2.upload(coefs2.data(), coefs2.size() * sizeof(FiberCoef), s)) !=
          hipSuccess ||
      (e = cp->wgg.upload(wgg.data(), wgg.size() * sizeof(int32_t), s)) != hipSuccess)
    return std::string("device table upload failed: ") + hipGetErrorString(e);
  return "";
}

hipError_t exec_chain(const ChainPlan& cp, int cplx, hipStream_t s, const void* X, const void* B1,
                      const void* B2, void* Z, std::string* err, KernelTimer* kt) {
  hipError_t e;
  {
    TimeScope ts(kt, "fiber_chain_kernel", s);
    e = launch_fiber_chain(s, cplx, cp.nf, cp.xm, cp.ym, (const ChainGroup*)cp.groups.p,
                           (const FiberSlot*)cp.slots.p, (const FiberOutElem*)cp.zouts.p,
                           (const uint32_t*)cp.ymask.p, (const FiberCoef*)cp.coefs1.p,
                           (const FiberCoef*)cp.coefs2.p, (const int32_t*)cp.wgg.p, cp.wgs, X, B1,
                           B2, Z);
  }
  if (e != hipSuccess) *err = std::string("fiber chain kernel launch: ") + hipGetErrorString(e);
  return e;
}

}  // namespace itg
