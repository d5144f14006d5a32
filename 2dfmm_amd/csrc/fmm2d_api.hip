// fmm2d_api.hip -- host side of the library: one context (what the reference
// keeps in the globals qtree/vors, quadtree.h:49-50), the parameter file
// reader (quadtree.c:12-27), the build sequence of make_tree
// (quadtree.c:102-171) on the GPU, and the C ABI of include/fmm2d.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fmm2d.h"
#include "fmm2d_internal.h"

namespace fmm2d {

static char g_err[1024];

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

// Grow-only device buffer.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t cap = 0;
  // contents are not kept
  int ensure(size_t n, hipStream_t s) {
    if (n <= cap) return 0;
    if (p) {
      FMM_HIP_TRY(hipStreamSynchronize(s));
      FMM_HIP_TRY(hipFree(p));
      p = nullptr;
      cap = 0;
    }
    size_t c = std::max<size_t>(n + n / 4, 64);
    FMM_HIP_TRY(hipMalloc(&p, c * sizeof(T)));
    cap = c;
    return 0;
  }
  // contents [0, keep) are kept
  int ensure_keep(size_t n, size_t keep, hipStream_t s) {
    if (n <= cap) return 0;
    size_t c = std::max<size_t>(n + n / 2, 64);
    T *q = nullptr;
    FMM_HIP_TRY(hipMalloc(&q, c * sizeof(T)));
    if (p && keep) FMM_HIP_TRY(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
    FMM_HIP_TRY(hipStreamSynchronize(s));
    if (p) FMM_HIP_TRY(hipFree(p));
    p = q;
    cap = c;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Context {
  // parameters (fmm_param.dat)
  int nterms = 0, num_p = 0;
  double theta = 0.0, r_min = 0.0;
  bool have_params = false;
  bool fixed_params = false;
  std::string param_path = "../cal_cond/fmm_param.dat";
  int mode = FMM2D_MODE_TREECODE;
  bool prefetch = true;
  hipStream_t stream = nullptr;
  // the evaluation's second stream (default priority) and its fork/join
  // events: the leaves after the front ones run there (launch_fmm_eval)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int nfront = 0;  // front leaves of the launch order of the current tree
  // the build's last kernels that only the evaluation reads (leaf-relative
  // sources, smooth coefficients) run on the side stream after ev_sfork;
  // ev_side marks their end, and side_pending says the main stream has not
  // waited for it yet (join_side)
  hipEvent_t ev_sfork = nullptr, ev_side = nullptr;
  // the core radii's largest and smallest, formed on the side stream during the tree build
  hipEvent_t ev_mm = nullptr;
  bool side_pending = false;
  bool init = false;
  // tree
  bool valid = false;
  int n = 0, nnodes = 0, depth = 0, nleaves = 0;
  int key_depth = kMaxDepth;
  int t_nterms = 0, t_num_p = 0, t_mode = 0;  // what the tree was built with
  double t_theta = 0.0;
  // accurate fast mode parameters and data
  int fmm_p = 16, fmm_leaf = 64;
  double fmm_theta = 0.5, fmm_cut = 4.0;
  double rcut = 0.0;
  int ntleaves = 0;  // non-empty leaves (targets of the fused kernel)
  DevBuf<float4> loc;
  DevBuf<double> smax;
  DevBuf<double4> loc64;
  DevBuf<int32_t> leaves, lat;
  DevBuf<uint32_t> lflag, lpos;
  // launch order of the target leaves (sort keys, then the order itself)
  DevBuf<uint64_t> lkey, lkey_alt;
  DevBuf<uint32_t> lord, lord_alt;
  DevBuf<double2> direct_ws;  // partial sums of the direct judge
  double smax_h = 0.0;        // largest core radius of the tree's blobs
  double wmax_h = 0.0;        // log2(e) / (smallest core radius)^2, capped
  double sw_h = 0.0;          // sqrt(wmax_h) when all core radii are equal, else 0
  // learned per-leaf parameters
  bool learned = false;
  float p_margin = 0.f, cut_margin = 0.f;
  // floor of the learned near radius (core radii): a far pair at c core radii
  // carries exp(-c^2) of its velocity as the neglected core term
  float cut_min = 2.5f;
  DevBuf<float> mlp_w, feat, mlp_out;
  DevBuf<double> mean_abs;
  DevBuf<double> leaf_abs;  // per node: sum of |cir| over a leaf's blobs (learned features)
  DevBuf<float2> leaf_par;   // learned (or explicitly set) per-leaf (p, cut)
  bool leaf_par_on = false;
  // mirror images: the tree holds n_real blobs and their images after them
  int n_real = -1;  // < 0: no images
  DevBuf<double> mxy, mcir, mcor, muv;
  // forests (batched systems)
  int nsys = 1, root_shift = -1;
  DevBuf<int32_t> doff;
  std::vector<int32_t> h_roots;
  // persistent evaluation and its graph
  DevBuf<uint32_t> ntl_dev, counter;
  int persist_blocks = 0;  // 0: one wave per leaf over a sized grid
  // splitting of target leaves with much near work (0: off)
  int split_budget = 8192;
  int far_burst = 128;  // far cells per MFMA burst (64..128)
  bool parity_leaves = true;  // treecode at the blobs: a wave per 64 blobs (else a lane per target)
  int l2l = 2;          // parent local expansions for parents of >= l2l target leaves (0: off)
  int start_cap = kFrontCap;  // entries per start list (0: children walk from the root)
  DevBuf<double2> plocal;
  DevBuf<int32_t> nodeli;
  DevBuf<uint32_t> pfront;  // start lists of the children's walks (kFrontCap per group)
  DevBuf<int32_t> pfcnt;
  DevBuf<SplitCk> ck;
  DevBuf<int32_t> nfirst;
  DevBuf<int2> leaf_ck;
  DevBuf<int32_t> split_list;  // the leaves with chunk records of the last evaluation
  DevBuf<double2> part;
  DevBuf<uint32_t> split_ctr;
  int ncu = 0;
  // smooth cells of the mixed form (one core radius): half side <= eta sigma,
  // at least smooth_min blobs; smooth_region_h = 2 eta sigma of the current
  // tree (0: none built)
  double smooth_eta = 0.5;
  int smooth_min = 48;
  double smooth_region_h = 0.0;
  int t_smooth_min = 0;  // smooth_min the current tree's coefficients used
  DevBuf<float4> scoef;
  DevBuf<uint8_t> sflag;
  DevBuf<int2> srange;
  DevBuf<uint32_t> sunits, sutmp;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  unsigned char graph_key[sizeof(FmmArgs) + 16] = {};
  DevBuf<double> xy, cir, cor;
  DevBuf<uint64_t> key, key_alt;
  DevBuf<uint32_t> idx, idx_alt, hist, hist_scan, perm, total;
  DevBuf<double> sx, sy, sg, ss, bbox_part, uv_dev, tmp_in, tmp_out;
  DevBuf<NodeRec> nodes;
  DevBuf<int32_t> gparent, split;
  DevBuf<int32_t> lt;         // level table of the tree build (see launch_tree_level)
  DevBuf<uint32_t> lstat;     // status words of the levels' split-list scans
  DevBuf<uint8_t> cflag;      // split flags of one level's children
  int32_t *h_lt = nullptr;    // pinned copy of the level table
  int depth_hint = 0;         // depth of the previous tree
  RootBox *h_root = nullptr;  // root square of the current single tree (pinned, read with the levels)
  DevBuf<double2> coef;
  DevBuf<RootBox> root;
  std::vector<int> split_off;
  // host side
  uint32_t *h_total = nullptr;  // pinned; [8]: chunk records taken by the last evaluation
  size_t ck_need = 0;
  double *h_uv = nullptr;
  size_t h_uv_cap = 0;
  hipEvent_t ev_uv = nullptr;
  bool uv_pending = false, uv_ready = false;
  std::vector<uint64_t> h_src;  // blob positions (bit patterns), caller order
  std::vector<int32_t> hash;
  size_t cursor = 0;
  // phase timing with events on the library stream:
  // 0 start, 1 keys, 2 sorted, 7 levels, 3 tree (leaf order, gather), 8 S2M/M2M,
  // 4 expansions (and the accurate modes' set-up) | 5, 9, 6 around the last
  // evaluation (9: after the fused pass)
  bool timing = false;
  hipEvent_t tev[10] = {};
  bool tev_rec[10] = {};  // recorded since timing was switched on
};

static Context g;
static std::mutex g_mu;

static int ctx_init() {
  if (g.init) return 0;
  int ndev = 0;
  FMM_HIP_TRY(hipGetDeviceCount(&ndev));
  if (ndev < 1) {
    set_error("no HIP device available");
    return -1;
  }
  FMM_HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
  FMM_HIP_TRY(hipStreamCreateWithFlags(&g.side, hipStreamNonBlocking));
  FMM_HIP_TRY(hipEventCreateWithFlags(&g.ev_fork, hipEventDisableTiming));
  FMM_HIP_TRY(hipEventCreateWithFlags(&g.ev_join, hipEventDisableTiming));
  FMM_HIP_TRY(hipEventCreateWithFlags(&g.ev_sfork, hipEventDisableTiming));
  FMM_HIP_TRY(hipEventCreateWithFlags(&g.ev_mm, hipEventDisableTiming));
  FMM_HIP_TRY(hipEventCreateWithFlags(&g.ev_side, hipEventDisableTiming));
  FMM_HIP_TRY(hipHostMalloc((void **)&g.h_total, 64, hipHostMallocDefault));
  // the level table, then the root square (read back together, pinned)
  FMM_HIP_TRY(hipHostMalloc((void **)&g.h_lt, sizeof(int32_t) * kLtWords + 64 + sizeof(RootBox),
                            hipHostMallocDefault));
  g.h_root = reinterpret_cast<RootBox *>(
      reinterpret_cast<char *>(g.h_lt) + ((sizeof(int32_t) * kLtWords + 63) & ~size_t(63)));
  memset(g.h_total, 0, 64);
  FMM_HIP_TRY(hipEventCreateWithFlags(&g.ev_uv, hipEventDisableTiming));
  for (auto &e : g.tev) FMM_HIP_TRY(hipEventCreate(&e));
  g.init = true;
  return 0;
}

// The main stream waits for the side stream's part of the last build.
static int join_side() {
  if (g.side_pending) {
    FMM_HIP_TRY(hipStreamWaitEvent(g.stream, g.ev_side, 0));
    g.side_pending = false;
  }
  return 0;
}

static inline void mark(int i) {
  if (g.timing && hipEventRecord(g.tev[i], g.stream) == hipSuccess) g.tev_rec[i] = true;
}

// quadtree.c:12-27 (the same fscanf formats, so "0.5d0" reads as 0.5 and the
// r_min line of the shipped file does not parse, which is harmless).
static int read_param_file() {
  FILE *f = fopen(g.param_path.c_str(), "r");
  if (!f) {
    printf("Error: Could not read FMM parameters.");
    fflush(stdout);
    if (g.have_params) return 0;
    set_error("cannot open parameter file %s and no parameters were set", g.param_path.c_str());
    return -1;
  }
  int nt = g.nterms, np = g.num_p;
  double th = g.theta, rm = g.r_min;
  int got = 0;
  if (fscanf(f, "%i    nterms", &nt) == 1) got++;
  if (fscanf(f, "%i   num_p", &np) == 1) got++;
  if (fscanf(f, "%lf theta", &th) == 1) got++;
  if (fscanf(f, "%lf  r_min ! for Tree", &rm) == 1) got++;
  fclose(f);
  if (got < 3) {
    if (g.have_params) return 0;
    set_error("parameter file %s: could not read nterms, num_p, theta", g.param_path.c_str());
    return -1;
  }
  if (nt < 1 || nt > kMaxTerms || np < 1 || !(th > 0.0)) {
    set_error("parameter file %s: nterms=%d (1..%d), num_p=%d (>=1), theta=%g (>0)",
              g.param_path.c_str(), nt, kMaxTerms, np, th);
    return -1;
  }
  g.nterms = nt;
  g.num_p = np;
  g.theta = th;
  g.r_min = rm;
  g.have_params = true;
  return 0;
}

// The parameter file holds the treecode's parameters; the accurate modes take
// theirs from fmm_set_fmm_params.
static int resolve_params() {
  if (g.fixed_params || g.mode != FMM2D_MODE_TREECODE) return 0;
  return read_param_file();
}

static int apply_learned();

// Tree levels the first key sort orders (two key bits per level, below the
// system bits): most trees are shallower, and a deeper one is sorted again on
// every bit.
constexpr int kSortLevels = 16;

// The tree build of make_tree on device inputs (caller order).  With nsys > 1
// the inputs are nsys independent systems, system s owning blobs
// [h_off[s], h_off[s+1]); each gets its own tree (a forest) in the same pass.
static int build(int n, const double *dxy, const double *dcir, const double *dcor, int nsys = 1,
                 const int32_t *h_off = nullptr) {
  hipStream_t s = g.stream;
  if (join_side()) return -1;  // the last build's side work reads buffers rebuilt here
  g.valid = false;
  g.uv_pending = g.uv_ready = false;
  g.hash.clear();
  g.cursor = 0;
  g.n = n;
  g.nsys = nsys;
  g.root_shift = -1;
  g.leaf_par_on = false;
  g.smooth_region_h = 0.0;
  g.n_real = -1;
  g.t_mode = g.mode;
  if (g.mode == FMM2D_MODE_TREECODE) {
    g.t_nterms = g.nterms;
    g.t_num_p = g.num_p;
    g.t_theta = g.theta;
  } else {
    g.t_nterms = g.fmm_p;
    g.t_num_p = g.fmm_leaf;
    g.t_theta = g.fmm_theta;
  }
  g.nnodes = g.depth = g.nleaves = g.ntleaves = 0;
  if (n == 0) {
    g.valid = true;
    return 0;
  }
  // accurate modes: the core radii's largest and smallest (near radius, pair
  // kernel scale, smooth cells) on the side stream, beside the sort and the
  // tree levels; the main stream waits for them before the launch keys
  if (g.t_mode != FMM2D_MODE_TREECODE) {
    if (g.smax.ensure(4, s)) return -1;
    FMM_HIP_TRY(hipEventRecord(g.ev_sfork, s));
    FMM_HIP_TRY(hipStreamWaitEvent(g.side, g.ev_sfork, 0));
    if (launch_minmax_reduce(dcor, n, g.smax.p, g.side)) return -1;
    FMM_HIP_TRY(hipEventRecord(g.ev_mm, g.side));
  }
  int sysbits = 0;
  while ((1 << sysbits) < nsys) sysbits++;
  const int D = nsys == 1 ? g.key_depth : std::min(g.key_depth, (64 - sysbits) / 2);
  const int nroot_slots = (nsys + 3) & ~3;
  if (g.key.ensure(n, s) || g.key_alt.ensure(n, s) || g.idx.ensure(n, s) || g.idx_alt.ensure(n, s) ||
      g.perm.ensure(n, s) || g.sx.ensure(n, s) || g.sy.ensure(n, s) || g.sg.ensure(n, s) ||
      g.ss.ensure(n, s) || g.root.ensure(nsys, s) || g.bbox_part.ensure(4 * 512, s) ||
      g.total.ensure(4, s) || g.hist.ensure(radix_sort_hist_words(n), s) ||
      g.hist_scan.ensure(radix_sort_hist_words(n), s))
    return -1;
  // node capacity: a first guess, doubled (and the build redone) whenever a
  // level launch reports that it ran out of room
  size_t est = std::max(g.nodes.cap, nroot_slots + 4 * ((size_t)n / (size_t)g.t_num_p + 1) + 64);
  int nroots = 1, child_base = 0;
  for (int full = 0;;) {
    if (g.nodes.ensure_keep(est, 0, s) || g.gparent.ensure_keep(g.nodes.cap / 4 + 1, 0, s) ||
        g.split.ensure_keep(g.nodes.cap / 4 + nsys + 1, 0, s))
      return -1;
    mark(0);
    if (nsys == 1) {
      if (launch_bbox(dxy, n, g.root.p, g.bbox_part.p, s)) return -1;
      if (launch_keygen(dxy, n, g.root.p, D, g.key.p, g.idx.p, s)) return -1;
      if (launch_init_root(g.nodes.p, g.gparent.p, g.root.p, n, g.split.p, s)) return -1;
    } else {
      g.root_shift = 2 * D;
      if (g.doff.ensure(nsys + 1, s)) return -1;
      FMM_HIP_TRY(hipMemcpyAsync(g.doff.p, h_off, sizeof(int32_t) * (nsys + 1),
                                 hipMemcpyHostToDevice, s));
      if (launch_forest_roots(dxy, n, g.doff.p, nsys, D, g.root.p, g.key.p, g.idx.p, g.nodes.p,
                              g.gparent.p, nroot_slots, s))
        return -1;
      // level 0 split list: the non-empty systems
      g.h_roots.clear();
      for (int k = 0; k < nsys; k++)
        if (h_off[k + 1] > h_off[k]) g.h_roots.push_back(k);
      nroots = (int)g.h_roots.size();
      if (nroots > 0)
        FMM_HIP_TRY(hipMemcpyAsync(g.split.p, g.h_roots.data(), sizeof(int32_t) * nroots,
                                   hipMemcpyHostToDevice, s));
    }
    mark(1);
    // the top sysbits + 2 kSortLevels key bits first: enough for a tree of
    // up to kSortLevels levels; a deeper tree is sorted again on every bit
    const int kbits = 2 * D + sysbits;
    const int lo_bit = full ? 0 : std::max(0, kbits - (sysbits + 2 * kSortLevels));
    const int lsort = lo_bit == 0 ? D : (kbits - lo_bit - sysbits) / 2;
    if (radix_sort_pairs(&g.key.p, &g.idx.p, &g.key_alt.p, &g.idx_alt.p, n, lo_bit, kbits,
                         g.hist.p, g.hist_scan.p, s))
      return -1;
    mark(2);
    // (the sort may swap key/key_alt and idx/idx_alt; each pair is always
    // sized together, so their capacities are equal and need no fix-up)
    // Levels 1..lsort, one launch each, with the counts kept on the device;
    // the host reads the level table back once per batch of launches (the
    // first batch reaches one level past the previous tree's depth).
    const int node_cap = (int)std::min<size_t>(g.nodes.cap, (size_t)INT32_MAX);
    const int split_cap = (int)std::min<size_t>(g.split.cap, (size_t)INT32_MAX);
    const int stat_len = tree_level_tiles(node_cap) + 2;
    if (g.lstat.ensure(2 * (size_t)stat_len, s) || g.lt.ensure(kLtWords, s) ||
        g.cflag.ensure((size_t)node_cap, s))
      return -1;
    if (launch_tree_levels_init(g.lt.p, nroots, nroot_slots, g.lstat.p, stat_len, s)) return -1;
    const int max_splits = (int)std::min<int64_t>((int64_t)n / std::max(1, g.t_num_p) + nsys,
                                                  (int64_t)split_cap);
    int launched = 0, depth = 0;
    bool more = false, overflow = false;
    for (;;) {
      const int upto = launched == 0 ? std::min(lsort, std::max(g.depth_hint + 1, 6)) : lsort;
      for (int level = launched + 1; level <= upto; level++) {
        // split nodes of level - 1: at most nroots 4^(level-1), and at most
        // one per num_p blobs
        int64_t bound = nroots;
        for (int k = 1; k < level && bound < max_splits; k++) bound *= 4;
        bound = std::min<int64_t>(bound, max_splits);
        if (launch_tree_level(g.key.p, 2 * D, level, D, g.t_num_p, (int)bound, g.nodes.p,
                              g.gparent.p, g.split.p, node_cap, split_cap, g.lt.p, g.cflag.p,
                              g.lstat.p, stat_len, s))
          return -1;
      }
      launched = upto;
      FMM_HIP_TRY(hipMemcpyAsync(g.h_lt, g.lt.p, sizeof(int32_t) * kLtWords, hipMemcpyDeviceToHost,
                                 s));
      if (nsys == 1)
        FMM_HIP_TRY(hipMemcpyAsync(g.h_root, g.root.p, sizeof(RootBox), hipMemcpyDeviceToHost, s));
      FMM_HIP_TRY(hipStreamSynchronize(s));
      const int err = g.h_lt[kLtErr];
      if (err & kLtErrDepth) {
        set_error("tree depth exceeds %d levels (are >= num_p=%d blobs coincident?)", D,
                  g.t_num_p);
        return -1;
      }
      if (err & kLtErrLookback) {
        set_error("tree build: a level's split-list scan did not complete");
        return -1;
      }
      if (err & kLtErrOverflow) {
        overflow = true;
        break;
      }
      if (g.h_lt[kLtS + launched] > 0 && launched < lsort) continue;  // deeper: next batch
      more = g.h_lt[kLtS + launched] > 0;  // deeper than the sorted bits reach
      depth = 1;
      while (depth < launched && g.h_lt[kLtS + depth] > 0) depth++;
      break;
    }
    if (overflow) {
      est = 2 * g.nodes.cap;
      continue;
    }
    if (more) {
      full = 1;
      continue;
    }
    g.depth = depth;
    g.depth_hint = depth;
    g.split_off.resize(depth + 2);
    for (int l = 0; l <= depth + 1; l++) g.split_off[l] = g.h_lt[kLtOff + l];
    child_base = g.h_lt[kLtCb + depth + 1];
    break;
  }
  mark(7);
  g.nnodes = child_base;
  g.nleaves = (child_base - nroot_slots + nroots) - g.split_off.back();
  if (launch_leaf_order(g.nodes.p, g.nnodes, g.idx.p, g.perm.p, g.t_num_p, s)) return -1;
  if (g.loc64.ensure(n, s) ||
      launch_gather(g.perm.p, n, dxy, dcir, dcor, g.loc64.p, g.sx.p, g.sy.p, g.sg.p, g.ss.p, s))
    return -1;
  mark(3);
  const size_t ncoef = (size_t)g.nnodes * g.t_nterms;
  if (g.coef.ensure(ncoef, s)) return -1;
  const bool accurate = g.t_mode != FMM2D_MODE_TREECODE;
  if (accurate && g.leaf_abs.ensure(g.nnodes, s)) return -1;
  if (launch_s2m(g.nodes.p, g.nnodes, nroot_slots, g.sx.p, g.sy.p, g.sg.p, g.t_nterms, g.coef.p,
                 accurate ? g.leaf_abs.p : nullptr, s))
    return -1;
  for (int l = g.depth - 1; l >= 1; l--) {
    const int a = g.split_off[l], b = g.split_off[l + 1];
    if (launch_m2m_level(g.nodes.p, g.split.p + a, b - a, g.t_nterms, g.coef.p, s)) return -1;
  }
  mark(8);
  if (g.t_mode != FMM2D_MODE_TREECODE) {
    // target leaves, leaf-relative FP32 sources, near-field radius
    if (g.leaves.ensure(n, s) || g.lat.ensure(n, s) || g.lflag.ensure(n, s) || g.lpos.ensure(n, s) ||
        g.loc.ensure(n, s) || g.loc64.ensure(n, s) || g.smax.ensure(4, s))
      return -1;
    // ntl_dev: [0] target leaves, [1] sibling groups (L2L), [2] front leaves
    if (g.ntl_dev.ensure(4, s) || g.counter.ensure(1, s)) return -1;
    if (launch_leaf_list(g.nodes.p, g.nnodes, n, g.lat.p, g.lflag.p, g.lpos.p, g.leaves.p,
                         g.ntl_dev.p, s))
      return -1;
    if (g.nfirst.ensure(g.nnodes, s) || g.nodeli.ensure(g.nnodes, s) ||
        g.plocal.ensure((size_t)(g.nnodes / 4 + 1) * 16, s) ||
        g.pfront.ensure((size_t)(g.nnodes / 4 + 1) * kFrontCap, s) ||
        g.pfcnt.ensure((size_t)(g.nnodes / 4 + 1), s))
      return -1;
    // smooth cells: mixed form with one core radius only (the radii are
    // checked on the device: the kernels below skip every node otherwise);
    // every internal node (>= num_p blobs) of the smooth level is a smooth cell
    const bool smooth_try = g.t_mode == FMM2D_MODE_MIXED && g.smooth_eta > 0.0;
    const int smin = std::min(g.smooth_min, g.t_num_p);
    if (smooth_try && (g.scoef.ensure((size_t)g.nnodes * kSmoothF4, s) ||
                       g.sflag.ensure(2 * (size_t)g.nnodes, s) || g.srange.ensure(g.nnodes, s) ||
                       g.sunits.ensure((size_t)g.nnodes + 1, s) || g.sutmp.ensure(g.nnodes, s)))
      return -1;
    // What only the fused pass reads runs on the side stream from here on,
    // beside the launch keys, the readback and the learned parameters: the
    // leaf-relative sources, the smooth coefficients and work units (their
    // scales come from the core radii's min/max, already formed on that
    // stream), then the launch order (event ev_side).  The evaluation waits
    // for it just before the fused pass; every other later call waits for it
    // at once (join_side).  The scans of the two streams share one workspace
    // (scan_u32): the main stream's last scan (the leaf list) is ordered
    // before the side stream's first.
    hipStream_t t = g.side;
    FMM_HIP_TRY(hipEventRecord(g.ev_sfork, s));
    FMM_HIP_TRY(hipStreamWaitEvent(t, g.ev_sfork, 0));
    if (launch_node_first_blob(g.nodes.p, g.nnodes, g.nfirst.p, t)) return -1;
    if (launch_leaf_local(g.nodes.p, g.leaves.p, g.ntl_dev.p, g.nleaves, g.sx.p, g.sy.p, g.sg.p,
                          g.ss.p, g.loc.p, g.loc64.p, g.smax.p, t))
      return -1;
    if (smooth_try && (launch_smooth_coef(g.nodes.p, 0, g.nnodes, g.nfirst.p, g.sx.p, g.sy.p,
                                          g.sg.p, g.smax.p, g.smooth_eta, smin, g.scoef.p, t) ||
                       launch_smooth_units(g.nodes.p, g.nnodes, g.smax.p, g.smooth_eta, smin,
                                           g.sutmp.p, g.sunits.p, t)))
      return -1;
    FMM_HIP_TRY(hipStreamWaitEvent(s, g.ev_mm, 0));
    // launch order keys (boxes large against the near radius first) and the
    // number of such front leaves, read back with the leaf count
    const int lbound = std::min(n, g.nnodes);
    if (g.lkey.ensure(lbound, s) || g.lkey_alt.ensure(lbound, s) || g.lord.ensure(lbound, s) ||
        g.lord_alt.ensure(lbound, s) ||
        launch_leaf_launch_keys(g.nodes.p, g.leaves.p, lbound, g.ntl_dev.p, g.smax.p, g.fmm_cut,
                                g.lkey.p, g.lord.p, g.ntl_dev.p + 2, s))
      return -1;
    double smm[4] = {0.0, 0.0, 0.0, 0.0};
    FMM_HIP_TRY(hipMemcpyAsync(g.h_total, g.ntl_dev.p, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               s));
    FMM_HIP_TRY(hipMemcpyAsync(smm, g.smax.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    FMM_HIP_TRY(hipStreamSynchronize(s));
    g.ntleaves = (int)g.h_total[0];
    g.nfront = (int)g.h_total[2];
    g.smax_h = smm[0];
    g.rcut = g.fmm_cut * smm[0];
    // the FP32 pair kernel's w = log2(e) / sigma^2 (capped at 7.2e19) and, with
    // one core radius, its square root (formed on the device)
    g.wmax_h = smm[2];
    g.sw_h = smm[3];
    const bool smooth = smooth_try && g.sw_h > 0.0;
    const double region = 2.0 * g.smooth_eta * smm[0];
    if (smooth) {
      g.smooth_region_h = region;
      g.t_smooth_min = smin;
    }
    // the learned parameters, then what the parent expansions read (node ->
    // target leaf, smooth runs) on the main stream, which the evaluation
    // continues
    g.valid = true;
    if (g.learned && apply_learned()) {
      g.valid = false;
      return -1;
    }
    if (launch_l2l_setup(g.leaves.p, g.ntleaves, g.nodeli.p, g.ntl_dev.p + 1, g.nnodes / 4, s))
      return -1;
    if (smooth && launch_smooth_ranges(g.nodes.p, g.gparent.p, g.nnodes, nroot_slots, region, smin,
                                       g.sflag.p, g.srange.p, s))
      return -1;
    // the launch order: one stable 8-bit radix pass over the keys
    if (radix_sort_pairs(&g.lkey.p, &g.lord.p, &g.lkey_alt.p, &g.lord_alt.p, g.ntleaves, 0, 8,
                         g.hist.p, g.hist_scan.p, t))
      return -1;
    FMM_HIP_TRY(hipEventRecord(g.ev_side, t));
    g.side_pending = true;
  }
  mark(4);
  g.valid = true;
  return 0;
}

static TreeView view() {
  TreeView tv;
  tv.nodes = g.nodes.p;
  tv.gparent = g.gparent.p;
  tv.coef = g.coef.p;
  tv.sx = g.sx.p;
  tv.sy = g.sy.p;
  tv.sg = g.sg.p;
  tv.ss = g.ss.p;
  tv.nterms = g.t_nterms;
  tv.theta = g.t_theta;
  tv.key = g.key.p;
  tv.root_shift = g.root_shift;
  return tv;
}

// join = false only for the evaluation, which waits for the side stream
// itself, after its parent expansions.
static int need_tree(bool join = true) {
  if (!g.init || !g.valid) {
    set_error("no valid tree: call make_tree first");
    return -1;
  }
  return join ? join_side() : 0;
}

static FmmArgs fmm_args(double *d_uv) {
  FmmArgs a;
  memset(&a, 0, sizeof a);  // padding too: the graph key compares the bytes
  a.nodes = g.nodes.p;
  a.gparent = g.gparent.p;
  a.coef = g.coef.p;
  a.sx = g.sx.p;
  a.sy = g.sy.p;
  a.sg = g.sg.p;
  a.loc64 = g.loc64.p;
  a.loc = g.loc.p;
  a.leaves = g.leaves.p;
  a.nleaves = g.ntleaves;
  a.lorder = g.t_mode != FMM2D_MODE_TREECODE ? g.lord.p : nullptr;
  a.pos0 = 0;
  a.npos = g.ntleaves;
  a.perm = g.perm.p;
  a.uv = reinterpret_cast<double2 *>(d_uv);
  a.p = g.t_nterms;
  a.theta = g.t_theta;
  a.rcut = g.rcut;
  a.key = g.key.p;
  a.root_shift = g.root_shift;
  a.leaf_par = g.leaf_par_on ? g.leaf_par.p : nullptr;
  a.smax = g.smax_h;
  a.wmax = (float)g.wmax_h;
  a.sw = g.sw_h;
  a.nfirst = g.nfirst.p;
  a.split_budget = g.split_budget;
  a.ck_cap = (int)(g.ck.cap / 2);  // two record regions (launch_fmm_eval)
  a.ck = g.ck.p;
  a.ck_count = g.split_ctr.p;
  a.leaf_ck = g.leaf_ck.p;
  a.split_list = g.split_list.p;
  a.split_count = g.split_ctr.p ? g.split_ctr.p + 4 : nullptr;
  a.part = g.part.p;
  a.smooth_region = g.smooth_region_h;
  a.smooth_min = g.t_smooth_min;
  a.scoef = g.scoef.p;
  a.srange = g.srange.p;
  a.sunits = g.sunits.p;
  a.far_burst = g.far_burst;
  a.l2l = (g.t_mode != FMM2D_MODE_TREECODE && g.plocal.p) ? g.l2l : 0;
  a.plocal = g.plocal.p;
  a.nodeli = g.nodeli.p;
  a.ngroups = g.ntl_dev.p ? g.ntl_dev.p + 1 : nullptr;
  a.ngroups_h = g.nnodes / 4;
  // start lists hold node indices below 2^27 (the rest of the word is a mask)
  a.pfront = (a.l2l && g.start_cap > 0 && g.nnodes < (1 << kSkipBits)) ? g.pfront.p : nullptr;
  a.pfcnt = g.pfcnt.p;
  a.front_cap = g.start_cap;
  return a;
}

// Buffers of the split passes for the current tree: room for two chunks per
// target leaf on average, or 1.25 x the most chunks an evaluation has asked
// for (a fused wave that finds no room runs its chunks itself, serially, so
// the buffer grows after an evaluation overflowed it), and the grid of the
// continuation passes.
static int ensure_split() {
  if (g.split_budget <= 0) return 0;
  hipStream_t s = g.stream;
  // chunks taken by the last evaluation (pinned copy, read without waiting:
  // a hint only)
  const size_t taken = std::max(g.h_total[8], g.h_total[10]);
  g.ck_need = std::max(g.ck_need, taken + taken / 4);
  const size_t cap = std::min(std::max(2 * (size_t)g.ntleaves + 4096, g.ck_need),
                              16 * (size_t)g.ntleaves + 4096);
  if (g.ck.ensure(2 * cap, s) || g.leaf_ck.ensure((size_t)g.ntleaves + 1, s) ||
      g.part.ensure(64 * g.ck.cap, s) || g.split_ctr.ensure(5, s) ||
      g.split_list.ensure((size_t)g.ntleaves + 1, s))
    return -1;
  if (g.ncu <= 0) {
    int dev = 0;
    FMM_HIP_TRY(hipGetDevice(&dev));
    FMM_HIP_TRY(hipDeviceGetAttribute(&g.ncu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  return 0;
}

static int split_blocks() { return 4 * (g.ncu > 0 ? g.ncu : 256); }

// Region 1 of the records (leaves after the front ones) rarely holds any in
// the mixed form, and a full persistent grid that finds none costs ~50 us;
// the full grid is used when the last evaluation wrote more than 64 there
// (h_total[10], read without waiting: a hint only).
static int side_blocks() { return g.h_total[10] > 64 ? split_blocks() : split_blocks() / 4; }

// Learned per-leaf (p, cut) for the current tree: features -> MLP -> params.
static int apply_learned() {
  hipStream_t s = g.stream;
  const int nl = g.ntleaves;
  if (nl <= 0) return 0;
  if (g.feat.ensure((size_t)nl * 8, s) || g.mlp_out.ensure((size_t)nl * 2, s) ||
      g.leaf_par.ensure(nl, s) || g.mean_abs.ensure(kMeanAbsWords, s))
    return -1;
  FmmArgs a = fmm_args(nullptr);
  if (launch_leaf_features(a, g.t_num_p, g.sg.p, g.n, g.leaf_abs.p, g.mean_abs.p, g.feat.p, s))
    return -1;
  if (launch_mlp_forward(g.mlp_w.p, nl, g.feat.p, g.mlp_out.p, s)) return -1;
  if (launch_leaf_params(g.mlp_out.p, nl, g.p_margin, g.cut_margin, g.cut_min, g.t_nterms,
                         (float)g.fmm_cut,
                         g.leaf_par.p, s))
    return -1;
  g.leaf_par_on = true;
  return 0;
}

// Persistent evaluation through a captured graph [reset counter, persistent
// kernel].  The launch reads its work size from device memory, so the graph
// stays valid across rebuilt trees; it is re-captured only when a pointer or
// parameter in the kernel arguments changes.
static int eval_fmm_graph(const FmmArgs &a, bool mixed) {
  unsigned char key[sizeof g.graph_key];
  memset(key, 0, sizeof key);
  FmmArgs ka = a;
  ka.nleaves = 0;    // the persistent kernel reads the count from the device
  ka.ngroups_h = 0;  // and so does the parent-expansion pass in this path
  memcpy(key, &ka, sizeof(FmmArgs));
  key[sizeof(FmmArgs)] = mixed ? 1 : 0;
  memcpy(key + sizeof(FmmArgs) + 4, &g.persist_blocks, sizeof(int));
  const int sb = split_blocks();
  memcpy(key + sizeof(FmmArgs) + 8, &sb, sizeof(int));
  if (!g.graph_exec || memcmp(key, g.graph_key, sizeof key) != 0) {
    if (g.graph_exec) (void)hipGraphExecDestroy(g.graph_exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.graph_exec = nullptr;
    g.graph = nullptr;
    FMM_HIP_TRY(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
    const int rc = launch_fmm_eval_persistent(a, mixed, g.ntl_dev.p, g.counter.p, g.persist_blocks,
                                              g.split_ctr.p, sb, g.stream);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(g.stream, &gr);
    if (rc || e != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      if (!rc) set_error("graph capture failed: %s", hipGetErrorString(e));
      return -1;
    }
    g.graph = gr;
    FMM_HIP_TRY(hipGraphInstantiate(&g.graph_exec, g.graph, nullptr, nullptr, 0));
    memcpy(g.graph_key, key, sizeof key);
  }
  FMM_HIP_TRY(hipGraphLaunch(g.graph_exec, g.stream));
  return 0;
}

static int eval_sources_dev(double *d_uv) {
  if (need_tree(false)) return -1;
  if (g.n == 0) return 0;
  if (g.n_real >= 0) {
    // mirror images ride along as sources: evaluate all, return the real blobs
    if (g.muv.ensure(2 * (size_t)g.n, g.stream)) return -1;
    const int nr = g.n_real;
    g.n_real = -1;
    const int rc = eval_sources_dev(g.muv.p);
    g.n_real = nr;
    if (rc) return rc;
    FMM_HIP_TRY(hipMemcpyAsync(d_uv, g.muv.p, sizeof(double) * 2 * nr, hipMemcpyDeviceToDevice,
                               g.stream));
    return 0;
  }
  mark(5);
  g.tev_rec[9] = false;  // set again only if this evaluation records it
  int rc;
  const bool mixed = g.t_mode == FMM2D_MODE_MIXED;
  if (g.t_mode != FMM2D_MODE_TREECODE && ensure_split()) return -1;
  if ((g.t_mode == FMM2D_MODE_TREECODE || g.persist_blocks > 0) && join_side()) return -1;
  // the default path waits for the build's side work in two parts: before
  // the parent expansions for what they read, before the fused pass for the
  // rest
  hipEvent_t ready1 = nullptr;  // what the parent expansions read is on the main stream
  hipEvent_t ready = g.side_pending ? g.ev_side : nullptr;
  if (g.t_mode == FMM2D_MODE_TREECODE)
    rc = g.parity_leaves ? launch_parity_leaves(view(), g.n, g.perm.p, d_uv, g.stream)
                         : launch_parity_eval(view(), g.n, nullptr, g.perm.p, d_uv, g.stream);
  else if (g.persist_blocks > 0 && g.use_graph)
    rc = eval_fmm_graph(fmm_args(d_uv), mixed);
  else if (g.persist_blocks > 0)
    rc = launch_fmm_eval_persistent(fmm_args(d_uv), mixed, g.ntl_dev.p, g.counter.p,
                                    g.persist_blocks, g.split_ctr.p, split_blocks(), g.stream);
  else
    rc = launch_fmm_eval(fmm_args(d_uv), mixed, g.ntl_dev.p, g.split_ctr.p, split_blocks(),
                         g.stream, g.timing ? g.tev[9] : nullptr, g.side, g.ev_fork, g.ev_join,
                         g.nfront, ready, side_blocks(), ready1);
  g.side_pending = false;
  if (!rc && g.timing && g.t_mode != FMM2D_MODE_TREECODE && g.persist_blocks == 0)
    g.tev_rec[9] = true;
  mark(6);
  if (!rc && g.t_mode != FMM2D_MODE_TREECODE && g.split_budget > 0)
    FMM_HIP_TRY(hipMemcpyAsync(g.h_total + 8, g.split_ctr.p, 3 * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, g.stream));
  return rc;
}

static int eval_batch_dev(int nt, const double *d_txy, double *d_uv) {
  if (need_tree()) return -1;
  if (g.nsys > 1) {
    set_error("a forest of systems evaluates at its blobs only (fmm_eval_sources)");
    return -1;
  }
  if (nt <= 0) return 0;
  if (g.n == 0) {
    FMM_HIP_TRY(hipMemsetAsync(d_uv, 0, sizeof(double) * 2 * (size_t)nt, g.stream));
    return 0;
  }
  if (g.t_mode != FMM2D_MODE_TREECODE) return launch_point_eval(fmm_args(nullptr), nt, d_txy, d_uv, g.stream);
  return launch_parity_eval(view(), nt, d_txy, nullptr, d_uv, g.stream);
}

// Pinned host buffer for the velocities at all blobs.
static int ensure_host_uv_buffer(size_t doubles) {
  if (doubles <= g.h_uv_cap) return 0;
  if (g.h_uv) (void)hipHostFree(g.h_uv);
  g.h_uv = nullptr;
  g.h_uv_cap = 0;
  if (hipHostMalloc((void **)&g.h_uv, sizeof(double) * doubles, hipHostMallocDefault) != hipSuccess) {
    set_error("pinned host allocation failed");
    return -1;
  }
  g.h_uv_cap = doubles;
  return 0;
}

// The velocities at all blobs (caller order) and the blob positions in host
// memory, for biot_tree's lookups: waits for make_tree's prefetch, or, when
// there was none (prefetch off, a tree built from device arrays), evaluates
// all blobs now -- in the accurate modes a blob's velocity is the all-blob
// evaluation's, whichever way the tree was built.  Positions of a tree built
// from device arrays come back from the leaf-ordered copies.
static int ensure_host_uv() {
  if (g.uv_pending) {
    FMM_HIP_TRY(hipEventSynchronize(g.ev_uv));
    g.uv_pending = false;
    g.uv_ready = true;
  }
  if (g.uv_ready) return 0;
  hipStream_t s = g.stream;
  const size_t nb = g.n_real >= 0 ? (size_t)g.n_real : (size_t)g.n;
  if (g.uv_dev.ensure(2 * (size_t)g.n, s) || ensure_host_uv_buffer(2 * nb)) return -1;
  if (eval_sources_dev(g.uv_dev.p)) return -1;
  FMM_HIP_TRY(hipMemcpyAsync(g.h_uv, g.uv_dev.p, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s));
  if (g.h_src.size() != 2 * nb) {
    std::vector<double> hx(g.n), hy(g.n);
    std::vector<uint32_t> hp(g.n);
    FMM_HIP_TRY(hipMemcpyAsync(hx.data(), g.sx.p, sizeof(double) * g.n, hipMemcpyDeviceToHost, s));
    FMM_HIP_TRY(hipMemcpyAsync(hy.data(), g.sy.p, sizeof(double) * g.n, hipMemcpyDeviceToHost, s));
    FMM_HIP_TRY(hipMemcpyAsync(hp.data(), g.perm.p, sizeof(uint32_t) * g.n, hipMemcpyDeviceToHost, s));
    FMM_HIP_TRY(hipStreamSynchronize(s));
    g.h_src.assign(2 * nb, 0);
    for (int j = 0; j < g.n; j++) {
      const uint32_t b = hp[j];
      if (b >= nb) continue;  // a mirror image, not a blob of the caller
      memcpy(&g.h_src[2 * (size_t)b], &hx[j], 8);
      memcpy(&g.h_src[2 * (size_t)b + 1], &hy[j], 8);
    }
    g.hash.clear();
    g.cursor = 0;
  }
  FMM_HIP_TRY(hipStreamSynchronize(s));
  g.uv_ready = true;
  return 0;
}

static int host_eval_batch(int nt, const double *xy, double *uv) {
  if (nt <= 0) return 0;
  hipStream_t s = g.stream;
  if (g.tmp_in.ensure(2 * (size_t)nt, s) || g.tmp_out.ensure(2 * (size_t)nt, s)) return -1;
  FMM_HIP_TRY(hipMemcpyAsync(g.tmp_in.p, xy, sizeof(double) * 2 * nt, hipMemcpyHostToDevice, s));
  if (eval_batch_dev(nt, g.tmp_in.p, g.tmp_out.p)) return -1;
  FMM_HIP_TRY(hipMemcpyAsync(uv, g.tmp_out.p, sizeof(double) * 2 * nt, hipMemcpyDeviceToHost, s));
  FMM_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

static inline uint64_t mix64(uint64_t a, uint64_t b) {
  uint64_t z = a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull + (a << 6) + (a >> 2));
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void build_hash() {
  const int nb = (int)(g.h_src.size() / 2);
  size_t cap = 16;
  while (cap < 2 * (size_t)nb) cap <<= 1;
  g.hash.assign(cap, -1);
  for (int i = 0; i < nb; i++) {
    size_t h = mix64(g.h_src[2 * i], g.h_src[2 * i + 1]) & (cap - 1);
    while (g.hash[h] >= 0) {
      const int j = g.hash[h];
      if (g.h_src[2 * j] == g.h_src[2 * i] && g.h_src[2 * j + 1] == g.h_src[2 * i + 1]) break;
      h = (h + 1) & (cap - 1);
    }
    if (g.hash[h] < 0) g.hash[h] = i;
  }
}

static int lookup(uint64_t bx, uint64_t by) {
  if (g.cursor < g.h_src.size() / 2 && g.h_src[2 * g.cursor] == bx && g.h_src[2 * g.cursor + 1] == by)
    return (int)g.cursor;
  if (g.hash.empty()) build_hash();
  const size_t cap = g.hash.size();
  size_t h = mix64(bx, by) & (cap - 1);
  while (g.hash[h] >= 0) {
    const int j = g.hash[h];
    if (g.h_src[2 * j] == bx && g.h_src[2 * j + 1] == by) return j;
    h = (h + 1) & (cap - 1);
  }
  return -1;
}

// Host nodes in gid order, with gids and levels.
struct HostTree {
  std::vector<NodeRec> nodes;
  std::vector<int32_t> gparent;
  std::vector<uint64_t> local, gid;
  std::vector<int32_t> level;
};

static int fetch_tree(HostTree &h) {
  hipStream_t s = g.stream;
  if (g.nsys > 1 || g.n_real >= 0) {
    set_error("tree dumps are for a single system without mirror images");
    return -1;
  }
  h.nodes.resize(g.nnodes);
  h.gparent.resize(g.nnodes / 4);
  FMM_HIP_TRY(hipMemcpyAsync(h.nodes.data(), g.nodes.p, sizeof(NodeRec) * g.nnodes,
                             hipMemcpyDeviceToHost, s));
  FMM_HIP_TRY(hipMemcpyAsync(h.gparent.data(), g.gparent.p, sizeof(int32_t) * (g.nnodes / 4),
                             hipMemcpyDeviceToHost, s));
  FMM_HIP_TRY(hipStreamSynchronize(s));
  h.local.assign(g.nnodes, 0);
  h.gid.assign(g.nnodes, 0);
  h.level.assign(g.nnodes, 0);
  for (int i = 4; i < g.nnodes; i++) {
    const int par = h.gparent[i >> 2];
    h.local[i] = (h.local[par] << 2) + (uint64_t)(i & 3);
    h.level[i] = h.level[par] + 1;
    h.gid[i] = h.local[i] + ((1ull << (2 * h.level[i])) - 1) / 3;
  }
  return 0;
}

}  // namespace fmm2d

using namespace fmm2d;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void make_tree(int nvor_b, const double *vor_b, const double *cir_b, const double *cor_b) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = ctx_init();
  if (!rc) rc = resolve_params();
  if (!rc && nvor_b < 0) {
    set_error("nvor_b < 0");
    rc = -1;
  }
  if (!rc) {
    const int n = nvor_b;
    hipStream_t s = g.stream;
    g.valid = false;
    if (n > 0) {
      if (g.xy.ensure(2 * (size_t)n, s) || g.cir.ensure(n, s) || g.cor.ensure(n, s)) rc = -1;
      if (!rc) {
        if (hipMemcpyAsync(g.xy.p, vor_b, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipMemcpyAsync(g.cir.p, cir_b, sizeof(double) * n, hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipMemcpyAsync(g.cor.p, cor_b, sizeof(double) * n, hipMemcpyHostToDevice, s) !=
                hipSuccess) {
          set_error("host to device copy of the blobs failed");
          rc = -1;
        }
      }
    }
    if (!rc) rc = build(n, g.xy.p, g.cir.p, g.cor.p);
    if (!rc) {
      g.h_src.resize(2 * (size_t)n);
      if (n > 0) memcpy(g.h_src.data(), vor_b, sizeof(double) * 2 * n);
      if (g.prefetch && n > 0) {
        if (g.uv_dev.ensure(2 * (size_t)n, s) || ensure_host_uv_buffer(2 * (size_t)n)) rc = -1;
        if (!rc) rc = eval_sources_dev(g.uv_dev.p);
        if (!rc && (hipMemcpyAsync(g.h_uv, g.uv_dev.p, sizeof(double) * 2 * n,
                                   hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipEventRecord(g.ev_uv, s) != hipSuccess)) {
          set_error("velocity download failed");
          rc = -1;
        }
        if (!rc) g.uv_pending = true;
      }
    }
  }
  if (rc) {
    g.valid = false;
    fprintf(stderr, "make_tree: %s\n", g_err);
  }
}

void biot_tree(double x, double y, double *u, double *v) {
  std::lock_guard<std::mutex> lk(g_mu);
  *u = 0.0;
  *v = 0.0;
  if (need_tree()) {
    fprintf(stderr, "biot_tree: %s\n", g_err);
    return;
  }
  if (g.n == 0) return;
  if (g.nsys > 1) {
    fprintf(stderr, "biot_tree: a forest of systems is evaluated by fmm_eval_sources\n");
    return;
  }
  if (g.t_mode != FMM2D_MODE_TREECODE) {
    // accurate modes: a blob's velocity is always the all-blob evaluation's
    if (ensure_host_uv()) {
      fprintf(stderr, "biot_tree: %s\n", g_err);
      return;
    }
  } else if (g.uv_pending) {
    if (hipEventSynchronize(g.ev_uv) != hipSuccess) {
      fprintf(stderr, "biot_tree: velocity prefetch failed\n");
      return;
    }
    g.uv_pending = false;
    g.uv_ready = true;
  }
  if (g.uv_ready) {
    uint64_t bx, by;
    memcpy(&bx, &x, 8);
    memcpy(&by, &y, 8);
    const int j = lookup(bx, by);
    if (j >= 0) {
      *u = g.h_uv[2 * j];
      *v = g.h_uv[2 * j + 1];
      g.cursor = (size_t)j + 1;
      return;
    }
  }
  double xy[2] = {x, y}, uv[2] = {0.0, 0.0};
  if (host_eval_batch(1, xy, uv)) {
    fprintf(stderr, "biot_tree: %s\n", g_err);
    return;
  }
  *u = uv[0];
  *v = uv[1];
}

int fmm_set_params(int nterms, int num_p, double theta, double r_min) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (nterms < 1 || nterms > kMaxTerms || num_p < 1 || !(theta > 0.0)) {
    set_error("fmm_set_params: nterms=%d (1..%d), num_p=%d (>=1), theta=%g (>0)", nterms,
              kMaxTerms, num_p, theta);
    return -1;
  }
  g.nterms = nterms;
  g.num_p = num_p;
  g.theta = theta;
  g.r_min = r_min;
  g.have_params = true;
  g.fixed_params = true;
  return 0;
}

int fmm_use_param_file(const char *path) {
  std::lock_guard<std::mutex> lk(g_mu);
  g.fixed_params = false;
  g.param_path = path ? path : "../cal_cond/fmm_param.dat";
  return 0;
}

int fmm_get_params(int *nterms, int *num_p, double *theta, double *r_min) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.have_params) {
    set_error("no parameters yet");
    return -1;
  }
  if (nterms) *nterms = g.nterms;
  if (num_p) *num_p = g.num_p;
  if (theta) *theta = g.theta;
  if (r_min) *r_min = g.r_min;
  return 0;
}

int fmm_set_mode(int mode) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (mode != FMM2D_MODE_TREECODE && mode != FMM2D_MODE_FMM64 && mode != FMM2D_MODE_MIXED) {
    set_error("fmm_set_mode: unknown mode %d", mode);
    return -1;
  }
  g.mode = mode;
  return 0;
}

int fmm_set_fmm_params(int p, int leaf, double theta, double cut_sigmas) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (p < 2 || p > 16 || leaf < 1 || leaf > 64 || !(theta > 0.0 && theta < 1.0) ||
      !(cut_sigmas >= 0.0)) {
    set_error("fmm_set_fmm_params: p=%d (2..16), leaf=%d (1..64), theta=%g (0..1), "
              "cut_sigmas=%g (>=0)", p, leaf, theta, cut_sigmas);
    return -1;
  }
  g.fmm_p = p;
  g.fmm_leaf = leaf;
  g.fmm_theta = theta;
  g.fmm_cut = cut_sigmas;
  return 0;
}

int fmm_selftest_mfma(int fp32, double *max_abs_err) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  double ha[16 * 8], hb[8 * 16], hout[256], href[256];
  for (int i = 0; i < 128; i++) {
    ha[i] = 0.25 * ((i * 37) % 17) - 2.0 + 0.001 * i;
    hb[i] = 0.5 * ((i * 11) % 13) - 3.0 - 0.002 * i;
  }
  DevBuf<double> d;
  if (d.ensure(128 + 128 + 256 + 256, g.stream)) return -1;
  int rc = 0;
  if (hipMemcpyAsync(d.p, ha, sizeof ha, hipMemcpyHostToDevice, g.stream) != hipSuccess ||
      hipMemcpyAsync(d.p + 128, hb, sizeof hb, hipMemcpyHostToDevice, g.stream) != hipSuccess)
    rc = -1;
  if (!rc) rc = launch_mfma_selftest(d.p, d.p + 128, d.p + 256, d.p + 512, fp32, g.stream);
  if (!rc && (hipMemcpyAsync(hout, d.p + 256, sizeof hout, hipMemcpyDeviceToHost, g.stream) !=
                  hipSuccess ||
              hipMemcpyAsync(href, d.p + 512, sizeof href, hipMemcpyDeviceToHost, g.stream) !=
                  hipSuccess ||
              hipStreamSynchronize(g.stream) != hipSuccess))
    rc = -1;
  d.release();
  if (rc) {
    set_error("mfma self-test failed to run");
    return -1;
  }
  double e = 0.0;
  for (int i = 0; i < 256; i++) e = std::max(e, std::fabs(hout[i] - href[i]));
  *max_abs_err = e;
  return 0;
}

int fmm_set_prefetch(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g.prefetch = on != 0;
  return 0;
}

const char *fmm_last_error(void) { return g_err; }

int fmm_make_tree_device(int n, const double *d_xy, const double *d_cir, const double *d_cor) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init() || resolve_params()) return -1;
  if (n < 0) {
    set_error("n < 0");
    return -1;
  }
  g.h_src.clear();
  return build(n, d_xy, d_cir, d_cor);
}

int fmm_make_tree_mirror_device(int n, const double *d_xy, const double *d_cir, const double *d_cor,
                                double d_mirror) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init() || resolve_params()) return -1;
  if (n < 0) {
    set_error("n < 0");
    return -1;
  }
  hipStream_t s = g.stream;
  const size_t m = 2 * (size_t)n;
  if (g.mxy.ensure(2 * m + 2, s) || g.mcir.ensure(m + 1, s) || g.mcor.ensure(m + 1, s)) return -1;
  if (n > 0) {
    FMM_HIP_TRY(hipMemcpyAsync(g.mxy.p, d_xy, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, s));
    FMM_HIP_TRY(hipMemcpyAsync(g.mcir.p, d_cir, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FMM_HIP_TRY(hipMemcpyAsync(g.mcor.p, d_cor, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (launch_mirror_images(n, d_xy, d_cir, d_mirror, g.mxy.p + 2 * (size_t)n, g.mcir.p + n,
                             g.mcor.p + n, s))
      return -1;
  }
  g.h_src.clear();
  if (build((int)m, g.mxy.p, g.mcir.p, g.mcor.p)) return -1;
  g.n_real = n;
  return 0;
}

int fmm_drift_ab3_device(int n, double *d_xy, const double *d_hist, double dt, int mirror,
                         double d_mirror) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  return launch_drift_ab3(n, d_xy, d_hist, dt, mirror, d_mirror, g.stream);
}

int fmm_history_shift_device(int n, double *d_hist, int first_step) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  return launch_history_shift(n, d_hist, first_step, g.stream);
}

int fmm_store_velocity_device(int n, const double *d_xy, const double *d_uv, double *d_hist,
                              double uinf, double vinf, double omz) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  return launch_store_velocity(n, d_xy, d_uv, d_hist, uinf, vinf, omz, g.stream);
}

int fmm_core_spread_device(int n, double *d_cor, double dt, double re) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  if (!(re > 0.0)) {
    set_error("Re must be > 0");
    return -1;
  }
  return launch_core_spread(n, d_cor, dt, re, g.stream);
}

static int check_offsets(int nsys, const int32_t *off) {
  if (nsys < 1 || !off || off[0] != 0) {
    set_error("forest: need nsys >= 1 and offsets starting at 0");
    return -1;
  }
  for (int k = 0; k < nsys; k++)
    if (off[k + 1] < off[k]) {
      set_error("forest: offsets must be non-decreasing");
      return -1;
    }
  return 0;
}

int fmm_make_forest_device(int nsys, const int32_t *offsets, const double *d_xy,
                           const double *d_cir, const double *d_cor) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init() || resolve_params() || check_offsets(nsys, offsets)) return -1;
  g.h_src.clear();
  return build(offsets[nsys], d_xy, d_cir, d_cor, nsys, offsets);
}

int fmm_make_forest(int nsys, const int32_t *offsets, const double *xy, const double *cir,
                    const double *cor) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init() || resolve_params() || check_offsets(nsys, offsets)) return -1;
  const int n = offsets[nsys];
  hipStream_t s = g.stream;
  g.valid = false;
  if (n > 0) {
    if (g.xy.ensure(2 * (size_t)n, s) || g.cir.ensure(n, s) || g.cor.ensure(n, s)) return -1;
    FMM_HIP_TRY(hipMemcpyAsync(g.xy.p, xy, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
    FMM_HIP_TRY(hipMemcpyAsync(g.cir.p, cir, sizeof(double) * n, hipMemcpyHostToDevice, s));
    FMM_HIP_TRY(hipMemcpyAsync(g.cor.p, cor, sizeof(double) * n, hipMemcpyHostToDevice, s));
  }
  g.h_src.clear();
  return build(n, g.xy.p, g.cir.p, g.cor.p, nsys, offsets);
}

int fmm_set_persistent(int blocks, int use_graph) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (blocks < 0) {
    set_error("blocks must be >= 0");
    return -1;
  }
  g.persist_blocks = blocks;
  g.use_graph = use_graph != 0 && blocks > 0;
  return 0;
}

int fmm_set_split(int budget) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (budget < 0) {
    set_error("split budget must be >= 0");
    return -1;
  }
  g.split_budget = budget;
  return 0;
}

int fmm_set_treecode_walk(int per_leaf) {
  std::lock_guard<std::mutex> lk(g_mu);
  g.parity_leaves = per_leaf != 0;
  return 0;
}

int fmm_set_far_burst(int cells) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (cells < 64 || cells > 128) {
    set_error("fmm_set_far_burst: %d cells (64..128)", cells);
    return -1;
  }
  g.far_burst = cells;
  return 0;
}

int fmm_set_l2l(int min_children) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (min_children < 0 || min_children > 4) {
    set_error("fmm_set_l2l: %d (0 = off, 1..4)", min_children);
    return -1;
  }
  g.l2l = min_children == 1 ? 2 : min_children;
  return 0;
}

int fmm_set_start_lists(int cap) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (cap < 0 || cap > kFrontCap) {
    set_error("fmm_set_start_lists: %d (0 = off, at most %d entries)", cap, kFrontCap);
    return -1;
  }
  g.start_cap = cap;
  return 0;
}

int fmm_get_split(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g.split_budget;
}

int fmm_set_smooth(double eta, int min_blobs) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!(eta >= 0.0) || eta > 0.6 || min_blobs < 1) {
    set_error("fmm_set_smooth: eta=%g (0 = off, at most 0.6), min_blobs=%d (>= 1)", eta,
              min_blobs);
    return -1;
  }
  g.smooth_eta = eta;
  g.smooth_min = min_blobs;
  return 0;
}

int fmm_get_smooth(double *eta, int *min_blobs) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (eta) *eta = g.smooth_eta;
  if (min_blobs) *min_blobs = g.smooth_min;
  return 0;
}

// ---- learned per-leaf parameters -------------------------------------------
int fmm_set_learned(const float *weights, float p_margin, float cut_margin) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  if (!weights) {
    g.learned = false;
    return 0;
  }
  if (g.mlp_w.ensure(kMlpWeights, g.stream)) return -1;
  FMM_HIP_TRY(hipMemcpyAsync(g.mlp_w.p, weights, sizeof(float) * kMlpWeights,
                             hipMemcpyHostToDevice, g.stream));
  FMM_HIP_TRY(hipStreamSynchronize(g.stream));
  g.learned = true;
  g.p_margin = p_margin;
  g.cut_margin = cut_margin;
  return 0;
}

int fmm_set_learned_floor(float cut_min) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!(cut_min >= 1.f && cut_min <= 6.f)) {
    set_error("fmm_set_learned_floor: cut_min=%g (1..6 core radii)", (double)cut_min);
    return -1;
  }
  g.cut_min = cut_min;
  return 0;
}

static int need_fmm_tree() {
  if (need_tree()) return -1;
  if (g.t_mode == FMM2D_MODE_TREECODE) {
    set_error("needs a tree built in an accurate mode");
    return -1;
  }
  return 0;
}

int fmm_num_target_leaves(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_fmm_tree()) return -1;
  return g.ntleaves;
}

int fmm_leaf_features(float *feat) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_fmm_tree()) return -1;
  const int nl = g.ntleaves;
  if (nl <= 0) return 0;
  hipStream_t s = g.stream;
  if (g.feat.ensure((size_t)nl * 8, s) || g.mean_abs.ensure(kMeanAbsWords, s)) return -1;
  if (launch_leaf_features(fmm_args(nullptr), g.t_num_p, g.sg.p, g.n, g.leaf_abs.p, g.mean_abs.p,
                           g.feat.p, s))
    return -1;
  FMM_HIP_TRY(hipMemcpyAsync(feat, g.feat.p, sizeof(float) * nl * 8, hipMemcpyDeviceToHost, s));
  FMM_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int fmm_blob_leaf(int32_t *leaf_of_blob) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_fmm_tree()) return -1;
  if (g.n == 0) return 0;
  hipStream_t s = g.stream;
  DevBuf<int32_t> d;
  if (d.ensure(g.n, s)) return -1;
  int rc = launch_blob_leaf(fmm_args(nullptr), d.p, s);
  // with mirror images only the real blobs (the first n_real) are reported
  const size_t nout = g.n_real >= 0 ? (size_t)g.n_real : (size_t)g.n;
  if (!rc && (hipMemcpyAsync(leaf_of_blob, d.p, sizeof(int32_t) * nout, hipMemcpyDeviceToHost, s) !=
                  hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess)) {
    set_error("download failed");
    rc = -1;
  }
  d.release();
  return rc;
}

int fmm_leaf_params(float *par) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_fmm_tree()) return -1;
  const int nl = g.ntleaves;
  if (!g.leaf_par_on) {
    for (int i = 0; i < nl; i++) {
      par[2 * i] = (float)g.t_nterms;
      par[2 * i + 1] = (float)g.fmm_cut;
    }
    return 0;
  }
  FMM_HIP_TRY(hipMemcpyAsync(par, g.leaf_par.p, sizeof(float) * 2 * nl, hipMemcpyDeviceToHost,
                             g.stream));
  FMM_HIP_TRY(hipStreamSynchronize(g.stream));
  return 0;
}

int fmm_set_leaf_params(const float *par, int n) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_fmm_tree()) return -1;
  if (!par) {
    g.leaf_par_on = false;
    return 0;
  }
  if (n != g.ntleaves) {
    set_error("fmm_set_leaf_params: %d entries for %d target leaves", n, g.ntleaves);
    return -1;
  }
  if (g.leaf_par.ensure(std::max(n, 1), g.stream)) return -1;
  FMM_HIP_TRY(hipMemcpyAsync(g.leaf_par.p, par, sizeof(float) * 2 * n, hipMemcpyHostToDevice,
                             g.stream));
  FMM_HIP_TRY(hipStreamSynchronize(g.stream));
  g.leaf_par_on = true;
  return 0;
}

int fmm_mlp_forward(int n, const float *x, const float *weights, float *out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  if (n <= 0) return 0;
  hipStream_t s = g.stream;
  DevBuf<float> dw, dx, dout;
  int rc = 0;
  if (dw.ensure(kMlpWeights, s) || dx.ensure((size_t)n * 8, s) || dout.ensure((size_t)n * 2, s))
    rc = -1;
  if (!rc && (hipMemcpyAsync(dw.p, weights, sizeof(float) * kMlpWeights, hipMemcpyHostToDevice, s) !=
                  hipSuccess ||
              hipMemcpyAsync(dx.p, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice, s) != hipSuccess))
    rc = -1;
  if (!rc) rc = launch_mlp_forward(dw.p, n, dx.p, dout.p, s);
  if (!rc && (hipMemcpyAsync(out, dout.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s) !=
                  hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess))
    rc = -1;
  (void)hipStreamSynchronize(s);
  dw.release(); dx.release(); dout.release();
  if (rc && !g_err[0]) set_error("mlp forward failed");
  return rc;
}

// gradient of mean((out - y)^2) over the n samples; *loss gets that mean
int fmm_mlp_grad(int n, const float *x, const float *y, const float *weights, float *grad,
                 float *loss) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  if (n <= 0) {
    set_error("n must be > 0");
    return -1;
  }
  hipStream_t s = g.stream;
  DevBuf<float> dw, dx, dy, scr, dg, dl;
  int rc = 0;
  if (dw.ensure(kMlpWeights, s) || dx.ensure((size_t)n * 8, s) || dy.ensure((size_t)n * 2, s) ||
      scr.ensure(mlp_scratch_floats(n), s) || dg.ensure(kMlpWeights, s) || dl.ensure(n, s))
    rc = -1;
  if (!rc && (hipMemcpyAsync(dw.p, weights, sizeof(float) * kMlpWeights, hipMemcpyHostToDevice, s) !=
                  hipSuccess ||
              hipMemcpyAsync(dx.p, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
              hipMemcpyAsync(dy.p, y, sizeof(float) * 2 * n, hipMemcpyHostToDevice, s) != hipSuccess))
    rc = -1;
  if (!rc) rc = launch_mlp_grad(dw.p, n, dx.p, dy.p, scr.p, dg.p, dl.p, s);
  std::vector<float> hl(n);
  if (!rc && (hipMemcpyAsync(grad, dg.p, sizeof(float) * kMlpWeights, hipMemcpyDeviceToHost, s) !=
                  hipSuccess ||
              hipMemcpyAsync(hl.data(), dl.p, sizeof(float) * n, hipMemcpyDeviceToHost, s) !=
                  hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess))
    rc = -1;
  (void)hipStreamSynchronize(s);
  dw.release(); dx.release(); dy.release(); scr.release(); dg.release(); dl.release();
  if (rc) {
    if (!g_err[0]) set_error("mlp gradient failed");
    return rc;
  }
  double l = 0.0;
  for (int i = 0; i < n; i++) l += hl[i];
  *loss = (float)l;
  return 0;
}

// Adam on minibatches of a fixed seeded shuffle per epoch; weights in/out;
// loss_hist[e] = mean loss of the minibatches of epoch e.
int fmm_mlp_train(int n, const float *x, const float *y, float *weights, int epochs, int batch,
                  float lr, unsigned seed, float *loss_hist) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  if (n <= 0 || epochs < 1 || batch < 1) {
    set_error("fmm_mlp_train: n, epochs, batch must be positive");
    return -1;
  }
  batch = std::min(batch, n);
  hipStream_t s = g.stream;
  DevBuf<float> dw, dx, dy, scr, dg, dm, dv, dl;
  int rc = 0;
  if (dw.ensure(kMlpWeights, s) || dx.ensure((size_t)n * 8, s) || dy.ensure((size_t)n * 2, s) ||
      scr.ensure(mlp_scratch_floats(batch), s) || dg.ensure(kMlpWeights, s) ||
      dm.ensure(kMlpWeights, s) || dv.ensure(kMlpWeights, s) || dl.ensure(batch, s))
    rc = -1;
  std::vector<float> hx((size_t)n * 8), hy((size_t)n * 2), hl(batch);
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  if (!rc && (hipMemcpyAsync(dw.p, weights, sizeof(float) * kMlpWeights, hipMemcpyHostToDevice, s) !=
                  hipSuccess ||
              hipMemsetAsync(dm.p, 0, sizeof(float) * kMlpWeights, s) != hipSuccess ||
              hipMemsetAsync(dv.p, 0, sizeof(float) * kMlpWeights, s) != hipSuccess))
    rc = -1;
  uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
  int t = 0;
  for (int e = 0; e < epochs && !rc; e++) {
    for (int i = n - 1; i > 0; i--) {  // Fisher-Yates with splitmix64
      st += 0x9E3779B97F4A7C15ull;
      uint64_t z = st;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      std::swap(order[i], order[(size_t)(z % (uint64_t)(i + 1))]);
    }
    for (int i = 0; i < n; i++) {
      memcpy(&hx[(size_t)i * 8], &x[(size_t)order[i] * 8], sizeof(float) * 8);
      memcpy(&hy[(size_t)i * 2], &y[(size_t)order[i] * 2], sizeof(float) * 2);
    }
    if (hipMemcpyAsync(dx.p, hx.data(), sizeof(float) * 8 * n, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(dy.p, hy.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, s) !=
            hipSuccess) {
      rc = -1;
      break;
    }
    double el = 0.0;
    int nb = 0;
    for (int b0 = 0; b0 + batch <= n && !rc; b0 += batch) {
      rc = launch_mlp_grad(dw.p, batch, dx.p + (size_t)b0 * 8, dy.p + (size_t)b0 * 2, scr.p, dg.p,
                           dl.p, s);
      if (!rc) rc = launch_adam(dw.p, dg.p, dm.p, dv.p, lr, ++t, s);
      if (!rc && (hipMemcpyAsync(hl.data(), dl.p, sizeof(float) * batch, hipMemcpyDeviceToHost, s) !=
                      hipSuccess ||
                  hipStreamSynchronize(s) != hipSuccess))
        rc = -1;
      if (!rc) {
        double l = 0.0;
        for (int k = 0; k < batch; k++) l += hl[k];
        el += l;
        nb++;
      }
    }
    if (loss_hist) loss_hist[e] = nb ? (float)(el / nb) : 0.f;
  }
  if (!rc && (hipMemcpyAsync(weights, dw.p, sizeof(float) * kMlpWeights, hipMemcpyDeviceToHost, s) !=
                  hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess))
    rc = -1;
  (void)hipStreamSynchronize(s);
  dw.release(); dx.release(); dy.release(); scr.release(); dg.release(); dm.release();
  dv.release(); dl.release();
  if (rc && !g_err[0]) set_error("mlp training failed");
  return rc;
}

int fmm_device_cus(void) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return -1;
  return cus;
}

int fmm_eval_sources_device(double *d_uv) {
  std::lock_guard<std::mutex> lk(g_mu);
  return eval_sources_dev(d_uv);
}

int fmm_eval_sources(double *uv) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_tree()) return -1;
  if (g.n == 0) return 0;
  hipStream_t s = g.stream;
  // with mirror images the tree holds 2 n_real blobs but only n_real are returned
  const size_t nout = g.n_real >= 0 ? (size_t)g.n_real : (size_t)g.n;
  if (g.uv_dev.ensure(2 * (size_t)g.n, s)) return -1;
  if (eval_sources_dev(g.uv_dev.p)) return -1;
  FMM_HIP_TRY(hipMemcpyAsync(uv, g.uv_dev.p, sizeof(double) * 2 * nout, hipMemcpyDeviceToHost, s));
  FMM_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int fmm_eval_batch(int nt, const double *xy, double *uv) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_tree()) return -1;
  return host_eval_batch(nt, xy, uv);
}

int fmm_eval_batch_device(int nt, const double *d_xy, double *d_uv) {
  std::lock_guard<std::mutex> lk(g_mu);
  return eval_batch_dev(nt, d_xy, d_uv);
}

int fmm_tree_size(int *nnodes, int *depth, int *nleaves) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_tree()) return -1;
  if (nnodes) *nnodes = g.n ? g.nnodes - 3 : 0;
  if (depth) *depth = g.depth;
  if (nleaves) *nleaves = g.nleaves;
  return 0;
}

int fmm_dump_tree(uint64_t *gid, int32_t *level, double *cx, double *cy, double *region,
                  int32_t *count, int32_t *is_leaf, int32_t *lo, int32_t *members, double *coeffs) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_tree()) return -1;
  if (g.n == 0) return 0;
  HostTree h;
  if (fetch_tree(h)) return -1;
  hipStream_t s = g.stream;
  std::vector<uint32_t> perm(g.n);
  std::vector<double2> co((size_t)g.nnodes * g.t_nterms);
  FMM_HIP_TRY(hipMemcpyAsync(perm.data(), g.perm.p, sizeof(uint32_t) * g.n, hipMemcpyDeviceToHost, s));
  FMM_HIP_TRY(hipMemcpyAsync(co.data(), g.coef.p, sizeof(double2) * co.size(), hipMemcpyDeviceToHost, s));
  FMM_HIP_TRY(hipStreamSynchronize(s));
  // node index order is level-major and Morton inside a level, i.e. ascending gid
  int k = 0, off = 0;
  for (int i = 0; i < g.nnodes; i++) {
    if (i >= 1 && i <= 3) continue;
    const NodeRec &r = h.nodes[i];
    const bool leaf = i != 0 && !node_internal(r.cnt);
    if (gid) gid[k] = h.gid[i];
    if (level) level[k] = h.level[i];
    if (cx) cx[k] = r.cx;
    if (cy) cy[k] = r.cy;
    if (region) region[k] = r.region;
    if (count) count[k] = node_count(r.cnt);
    if (is_leaf) is_leaf[k] = leaf ? 1 : 0;
    if (lo) lo[k] = off;
    if (leaf) {
      for (int j = 0; j < node_count(r.cnt); j++)
        if (members) members[off + j] = (int32_t)perm[r.first + j];
      off += node_count(r.cnt);
    }
    if (coeffs)
      for (int j = 0; j < g.t_nterms; j++) {
        coeffs[((size_t)k * g.t_nterms + j) * 2] = co[(size_t)i * g.t_nterms + j].x;
        coeffs[((size_t)k * g.t_nterms + j) * 2 + 1] = co[(size_t)i * g.t_nterms + j].y;
      }
    k++;
  }
  if (lo) lo[k] = off;
  return 0;
}

int fmm_dump_lists(int nt, const double *xy, int cap, uint64_t *gid, int32_t *kind, int32_t *len) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_tree()) return -1;
  if (nt <= 0) return 0;
  if (cap < 1) {
    set_error("cap must be >= 1");
    return -1;
  }
  if (g.n == 0) {
    for (int t = 0; t < nt; t++) len[t] = 0;
    return 0;
  }
  hipStream_t s = g.stream;
  HostTree h;
  if (fetch_tree(h)) return -1;
  DevBuf<double> dxy;
  DevBuf<int32_t> dl, dn;
  int rc = 0;
  if (dxy.ensure(2 * (size_t)nt, s) || dl.ensure((size_t)nt * cap * 2, s) || dn.ensure(nt, s)) rc = -1;
  std::vector<int32_t> hl;
  if (!rc) {
    if (hipMemcpyAsync(dxy.p, xy, sizeof(double) * 2 * nt, hipMemcpyHostToDevice, s) != hipSuccess) {
      set_error("copy failed");
      rc = -1;
    }
  }
  if (!rc) rc = launch_parity_lists(view(), nt, dxy.p, cap, dl.p, dn.p, s);
  if (!rc) {
    hl.resize((size_t)nt * cap * 2);
    if (hipMemcpyAsync(hl.data(), dl.p, sizeof(int32_t) * hl.size(), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(len, dn.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("list download failed");
      rc = -1;
    }
  }
  if (!rc) {
    for (int t = 0; t < nt; t++) {
      const int m = std::min(len[t], cap);
      for (int i = 0; i < m; i++) {
        const size_t o = (size_t)t * cap + i;
        gid[o] = h.gid[hl[2 * o]];
        kind[o] = hl[2 * o + 1];
      }
    }
  }
  (void)hipStreamSynchronize(s);
  dxy.release();
  dl.release();
  dn.release();
  return rc;
}

int fmm_direct_device(int ns, const double *d_xy, const double *d_cir, const double *d_cor, int nt,
                      const double *d_txy, double *d_uv, int style) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  if (style != 0 && style != 1) {
    set_error("style must be 0 or 1");
    return -1;
  }
  if (nt <= 0) return 0;
  const size_t need = (size_t)direct_chunks(ns, nt) * nt;
  if (g.direct_ws.ensure(need, g.stream)) return -1;
  return launch_direct(ns, d_xy, d_cir, d_cor, nt, d_txy, d_uv, style, g.direct_ws.p,
                       g.direct_ws.cap, g.stream);
}

int fmm_direct(int ns, const double *xy, const double *cir, const double *cor, int nt,
               const double *txy, double *uv, int style) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  if (style != 0 && style != 1) {
    set_error("style must be 0 or 1");
    return -1;
  }
  if (nt <= 0) return 0;
  hipStream_t s = g.stream;
  DevBuf<double> a, b, c, t, o;
  int rc = 0;
  const size_t nsz = ns > 0 ? (size_t)ns : 1;
  if (a.ensure(2 * nsz, s) || b.ensure(nsz, s) || c.ensure(nsz, s) || t.ensure(2 * (size_t)nt, s) ||
      o.ensure(2 * (size_t)nt, s))
    rc = -1;
  if (!rc && ns > 0 &&
      (hipMemcpyAsync(a.p, xy, sizeof(double) * 2 * ns, hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(b.p, cir, sizeof(double) * ns, hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(c.p, cor, sizeof(double) * ns, hipMemcpyHostToDevice, s) != hipSuccess)) {
    set_error("copy failed");
    rc = -1;
  }
  if (!rc && hipMemcpyAsync(t.p, txy, sizeof(double) * 2 * nt, hipMemcpyHostToDevice, s) != hipSuccess) {
    set_error("copy failed");
    rc = -1;
  }
  if (!rc && g.direct_ws.ensure((size_t)direct_chunks(ns, nt) * nt, s)) rc = -1;
  if (!rc)
    rc = launch_direct(ns, a.p, b.p, c.p, nt, t.p, o.p, style, g.direct_ws.p, g.direct_ws.cap, s);
  if (!rc && (hipMemcpyAsync(uv, o.p, sizeof(double) * 2 * nt, hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess)) {
    set_error("copy failed");
    rc = -1;
  }
  (void)hipStreamSynchronize(s);
  a.release(); b.release(); c.release(); t.release(); o.release();
  return rc;
}

int fmm_set_timing(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx_init()) return -1;
  g.timing = on != 0;
  for (bool &r : g.tev_rec) r = false;
  return 0;
}

int fmm_get_timing(double *ms, int n) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.init || !g.timing) {
    set_error("timing is off");
    return -1;
  }
  FMM_HIP_TRY(hipStreamSynchronize(g.stream));
  const int pairs[10][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {5, 6},
                            {2, 7}, {7, 3}, {3, 8}, {5, 9}, {9, 6}};
  for (int i = 0; i < n && i < 10; i++) {
    float t = -1.f;
    // only events recorded since timing was switched on are queried, so no
    // error is left behind for a later launch check to report
    if (g.tev_rec[pairs[i][0]] && g.tev_rec[pairs[i][1]] &&
        hipEventElapsedTime(&t, g.tev[pairs[i][0]], g.tev[pairs[i][1]]) != hipSuccess) {
      (void)hipGetLastError();
      t = -1.f;
    }
    ms[i] = t;
  }
  return 0;
}

// Work of the evaluation with all blobs as targets: out[0] far cells (M2L
// contractions of target leaves; treecode: M2S of (target, cell)), [1] near
// pairs, [2] (target, smooth cell) evaluations, [3] sum over target leaves
// of far cells x terms^2, [4] sum of targets x terms (L2P), [5] target leaves,
// [6] far cells of parent expansions (L2L), [7] their sum of far cells x the
// parent's terms^2, [8] parent expansions, [9] L2L shifts (target leaves that
// inherit one).
constexpr int kWorkOut = 10;
static int count_work(long long *out) {
  if (need_tree()) return -1;
  for (int i = 0; i < kWorkOut; i++) out[i] = 0;
  if (g.n == 0) return 0;
  if (g.t_mode != FMM2D_MODE_TREECODE) {
    DevBuf<int64_t> c;
    const size_t nw = (size_t)kCountWords * std::max(g.ntleaves, 1);
    if (c.ensure(nw, g.stream)) return -1;
    std::vector<int64_t> h(nw);
    int rc = launch_fmm_count(fmm_args(nullptr), c.p, g.stream);
    if (!rc && (hipMemcpyAsync(h.data(), c.p, sizeof(int64_t) * (size_t)kCountWords * g.ntleaves,
                               hipMemcpyDeviceToHost, g.stream) != hipSuccess ||
                hipStreamSynchronize(g.stream) != hipSuccess)) {
      set_error("count download failed");
      rc = -1;
    }
    c.release();
    if (rc) return rc;
    for (int i = 0; i < g.ntleaves; i++) {
      const int64_t *o = &h[(size_t)kCountWords * i];
      out[0] += o[0];
      out[1] += o[1];
      out[2] += o[2];
      out[3] += o[0] * o[3] * o[3];
      out[4] += o[4] * o[3];
    }
    out[5] = g.ntleaves;
    const FmmArgs a = fmm_args(nullptr);
    if (a.l2l) {
      const size_t ng = (size_t)(g.nnodes / 4);
      DevBuf<int64_t> pc;
      std::vector<int64_t> hp(3 * ng);
      if (pc.ensure(3 * ng + 3, g.stream)) return -1;
      rc = launch_parent_local(a, g.t_mode == FMM2D_MODE_MIXED, pc.p, split_blocks(), g.stream);
      if (!rc && (hipMemcpyAsync(hp.data(), pc.p, sizeof(int64_t) * 3 * ng, hipMemcpyDeviceToHost,
                                 g.stream) != hipSuccess ||
                  hipStreamSynchronize(g.stream) != hipSuccess)) {
        set_error("count download failed");
        rc = -1;
      }
      pc.release();
      if (rc) return rc;
      for (size_t k = 0; k < ng; k++) {
        out[6] += hp[3 * k];
        out[7] += hp[3 * k] * hp[3 * k + 1] * hp[3 * k + 1];
        out[8] += hp[3 * k + 2] > 0 ? 1 : 0;
        out[9] += hp[3 * k + 2];
      }
    }
    return 0;
  }
  DevBuf<int32_t> c;
  if (c.ensure(2 * (size_t)g.n, g.stream)) return -1;
  std::vector<int32_t> h(2 * (size_t)g.n);
  int rc = launch_parity_counts(view(), g.n, nullptr, c.p, g.stream);
  if (!rc && (hipMemcpyAsync(h.data(), c.p, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost,
                             g.stream) != hipSuccess ||
              hipStreamSynchronize(g.stream) != hipSuccess)) {
    set_error("count download failed");
    rc = -1;
  }
  c.release();
  if (rc) return rc;
  for (int i = 0; i < g.n; i++) {
    out[0] += h[2 * i];
    out[1] += h[2 * i + 1];
  }
  return 0;
}

int fmm_count_interactions(long long *m2s, long long *pairs) {
  std::lock_guard<std::mutex> lk(g_mu);
  long long o[kWorkOut];
  if (count_work(o)) return -1;
  *m2s = o[0];
  *pairs = o[1];
  return 0;
}

int fmm_count_work(long long *m2s, long long *pairs, long long *smooth) {
  std::lock_guard<std::mutex> lk(g_mu);
  long long o[kWorkOut];
  if (count_work(o)) return -1;
  *m2s = o[0];
  *pairs = o[1];
  *smooth = o[2];
  return 0;
}

int fmm_count_work_leaves(long long *out, int nleaves) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (need_fmm_tree()) return -1;
  if (nleaves != g.ntleaves) {
    set_error("fmm_count_work_leaves: %d entries for %d target leaves", nleaves, g.ntleaves);
    return -1;
  }
  if (nleaves == 0) return 0;
  DevBuf<int64_t> c;
  const size_t nw = (size_t)kCountWords * nleaves;
  if (c.ensure(nw, g.stream)) return -1;
  int rc = launch_fmm_count(fmm_args(nullptr), c.p, g.stream);
  if (!rc && (hipMemcpyAsync(out, c.p, sizeof(int64_t) * nw, hipMemcpyDeviceToHost, g.stream) !=
                  hipSuccess ||
              hipStreamSynchronize(g.stream) != hipSuccess)) {
    set_error("count download failed");
    rc = -1;
  }
  c.release();
  return rc;
}

int fmm_count_work_detail(long long *out, int n) {
  std::lock_guard<std::mutex> lk(g_mu);
  long long o[kWorkOut];
  if (count_work(o)) return -1;
  for (int i = 0; i < n && i < kWorkOut; i++) out[i] = o[i];
  return 0;
}

int fmm_synchronize(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.init) return 0;
  if (join_side()) return -1;
  FMM_HIP_TRY(hipStreamSynchronize(g.stream));
  return 0;
}

void fmm_release(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.init) return;
  (void)hipStreamSynchronize(g.side);
  (void)hipStreamSynchronize(g.stream);
  g.side_pending = false;
  g.xy.release(); g.cir.release(); g.cor.release();
  g.key.release(); g.key_alt.release(); g.idx.release(); g.idx_alt.release();
  g.hist.release(); g.hist_scan.release(); g.perm.release(); g.lt.release(); g.lstat.release(); g.cflag.release();
  g.total.release(); g.sx.release(); g.sy.release(); g.sg.release(); g.ss.release();
  g.bbox_part.release(); g.uv_dev.release(); g.tmp_in.release(); g.tmp_out.release();
  g.nodes.release(); g.gparent.release(); g.split.release(); g.coef.release(); g.root.release();
  g.loc.release(); g.loc64.release(); g.smax.release(); g.leaves.release(); g.lat.release();
  g.lflag.release(); g.lpos.release(); g.lkey.release(); g.lkey_alt.release(); g.lord.release();
  g.lord_alt.release(); g.direct_ws.release();
  g.doff.release(); g.ntl_dev.release(); g.counter.release();
  g.ck.release(); g.leaf_ck.release(); g.split_list.release(); g.part.release(); g.split_ctr.release();
  g.nfirst.release();
  g.plocal.release();
  g.nodeli.release();
  g.pfront.release();
  g.pfcnt.release();
  g.scoef.release();
  g.sflag.release();
  g.srange.release();
  g.sunits.release();
  g.sutmp.release();
  if (g.graph_exec) (void)hipGraphExecDestroy(g.graph_exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.graph_exec = nullptr;
  g.graph = nullptr;
  if (g.h_uv) (void)hipHostFree(g.h_uv);
  g.h_uv = nullptr;
  g.h_uv_cap = 0;
  g.valid = false;
}

}  // extern "C"
