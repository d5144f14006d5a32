// fmm2d_internal.h -- shared types of the HIP quadtree / treecode library.
//
// Device data layout (HBM), for N blobs and M tree nodes:
//   xy[N][2], cir[N], cor[N]          caller's order (input, read once)
//   key[N] u64, idx[N] u32            Morton path keys + blob ids (radix sorted)
//   perm[N] u32                       sorted position -> blob id (leaf order,
//                                     ascending blob id inside each leaf)
//   sx, sy, sg, ss [N] f64            blobs gathered into leaf order (SoA)
//   nodes[M] NodeRec (32 B)           level by level; the four children of a
//                                     node are consecutive, first index % 4 == 0
//   gparent[M/4] i32                  parent of each group of four siblings
//   coef[M][P] double2                multipole coefficients a_k, k < nterms
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>

namespace fmm2d {

// Maximum tree depth: keys hold 2 bits per level in a u64.
constexpr int kMaxDepth = 31;
// Largest supported expansion length (nterms).
constexpr int kMaxTerms = 32;

// One tree node.  `first` is the index of child q=0 for an internal node and
// the start of the node's blobs in the leaf-ordered arrays for a leaf; `cnt`
// is the blob count with kInternalBit set for internal (split) nodes.
struct __align__(32) NodeRec {
  double cx, cy, region;
  int32_t first;
  int32_t cnt;
};
constexpr int32_t kInternalBit = int32_t(0x80000000u);
__host__ __device__ inline int32_t node_count(int32_t c) { return c & 0x7fffffff; }
__host__ __device__ inline bool node_internal(int32_t c) { return (c & kInternalBit) != 0; }

// Root square, computed on the device (quadtree.c:81-99).
struct RootBox {
  double cx, cy, region;
  double pad;
};

#define FMM_HIP_TRY(expr)                                                         \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::fmm2d::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #expr,              \
                         hipGetErrorString(e_));                                  \
      return -1;                                                                  \
    }                                                                             \
  } while (0)

void set_error(const char *fmt, ...);

// ---- tree construction (tree_kernels.hip) --------------------------------
// Exclusive scan of n u32 values; total written to *d_total when non-null.
int scan_u32(const uint32_t *d_in, uint32_t *d_out, int n, uint32_t *d_total, hipStream_t s);
int launch_bbox(const double *d_xy, int n, RootBox *d_root, double *d_partial, hipStream_t s);
int launch_keygen(const double *d_xy, int n, const RootBox *d_root, int depth, uint64_t *d_key,
                  uint32_t *d_idx, hipStream_t s);
// Stable LSD radix sort of (key, idx) pairs on bits [lo_bit, bits) (8-bit
// digits from lo_bit up; pairs equal on those bits keep their input order).
// Uses the alternate buffers; the sorted result ends in (*key, *idx)
// (pointers swapped).
int radix_sort_pairs(uint64_t **key, uint32_t **idx, uint64_t **key_alt, uint32_t **idx_alt,
                     int n, int lo_bit, int bits, uint32_t *d_hist, uint32_t *d_hist_scan,
                     hipStream_t s);
size_t radix_sort_hist_words(int n);
// Forest of nsys systems (blobs [off[s], off[s+1]) on the device): per-system
// root squares, keys (s << 2*depth) | path, root nodes 0..nsys-1 and padding
// up to nroot_slots (a multiple of 4).
int launch_forest_roots(const double *d_xy, int n, const int32_t *d_off, int nsys, int depth,
                        RootBox *d_roots, uint64_t *d_key, uint32_t *d_idx, NodeRec *d_nodes,
                        int32_t *d_gparent, int nroot_slots, hipStream_t s);
int launch_init_root(NodeRec *d_nodes, int32_t *d_gparent, const RootBox *d_root, int n,
                     int32_t *d_split, hipStream_t s);
// Level table of the tree build, on the device (int32): per level l the
// number of split nodes S[l], the index of level l's first node CB[l], the
// start of level l's split nodes in the split list OFF[l], the ticket counter
// of level l's launch; and error flags.
constexpr int kLtLevels = 40;
constexpr int kLtS = 0, kLtCb = kLtLevels, kLtOff = 2 * kLtLevels, kLtTicket = 3 * kLtLevels,
              kLtErr = 4 * kLtLevels, kLtWords = 4 * kLtLevels + 1;
constexpr int kLtErrOverflow = 1, kLtErrDepth = 2, kLtErrLookback = 4;
static_assert(kLtLevels >= kMaxDepth + 3, "level table");
// Clears the level table and the two status arrays (stat_len words each); level
// 0 = the nroots roots already in the split list, level 1 starts at nroot_slots.
int launch_tree_levels_init(int32_t *d_lt, int nroots, int nroot_slots, uint32_t *d_stat,
                            int stat_len, hipStream_t s);
// Children (level `level`) of the split nodes of level-1, from the device
// level table, and the next split list; split_bound >= the number of those
// split nodes (sizes the grids).  Capacities: node_cap nodes (and d_cflag
// bytes), split_cap split-list entries, stat_len compaction tiles
// (tree_level_tiles of the children); past them the overflow flag is set.
int launch_tree_level(const uint64_t *d_key, int depth_bits, int level, int max_level, int num_p,
                      int split_bound, NodeRec *d_nodes, int32_t *d_gparent, int32_t *d_split,
                      int node_cap, int split_cap, int32_t *d_lt, uint8_t *d_cflag,
                      uint32_t *d_stat, int stat_len, hipStream_t s);
int tree_level_tiles(int nchild);
int launch_leaf_order(const NodeRec *d_nodes, int nnodes, const uint32_t *d_idx_sorted,
                      uint32_t *d_perm, int num_p, hipStream_t s);
// Blobs into leaf order (d_rec: n records of scratch).
int launch_gather(const uint32_t *d_perm, int n, const double *d_xy, const double *d_cir,
                  const double *d_cor, double4 *d_rec, double *sx, double *sy, double *sg,
                  double *ss, hipStream_t s);

// ---- expansions (expansion_kernels.hip) ----------------------------------
// S2M of every non-empty leaf; zeros for empty leaves and the first
// nroot_slots nodes (roots); internal nodes are left to M2M.  d_leaf_abs
// (optional, per node): the sum of |cir| over each leaf's blobs.
int launch_s2m(const NodeRec *d_nodes, int nnodes, int nroot_slots, const double *sx,
               const double *sy, const double *sg, int nterms, double2 *d_coef, double *d_leaf_abs,
               hipStream_t s);
int launch_m2m_level(const NodeRec *d_nodes, const int32_t *d_parents, int nparents, int nterms,
                     double2 *d_coef, hipStream_t s);

// Smooth cells (mixed form, one core radius sigma): a cell small against the
// core (half side <= eta sigma) is summed for every target through a Taylor
// expansion of the regularised kernel f(z) = (1 - e^{-|z|^2}) / z (z in core
// radii) about its centre, of total degree kSmoothQ in (d, conj d).  Per
// cell kSmoothC complex FP32 coefficients: nu_q .. nu_0 (the n = 0 column),
// then d_ab of the n >= 1 part in real monomials x^a y^b, in Horner order
// (b = q-1 .. 0, a = q-1-b .. 0); see smooth_coef_kernel.  Coordinates of the evaluation are the FP32
// leaf-relative ones scaled by sw = sqrt(log2 e) / sigma.
constexpr int kSmoothQ = 12;
constexpr int kSmoothC = (kSmoothQ + 1) + kSmoothQ * (kSmoothQ + 1) / 2;
constexpr int kSmoothF4 = (kSmoothC + 1) / 2;  // float4 per cell
// Coefficients of every node in [node0, node1) of region in (region/2,
// region], region = 2 eta sigma, with at least smin blobs (those of node k are
// sx..sg[nfirst[k] ...]), at d_out[node * kSmoothF4]; other nodes are left as
// they are, and so is every node without one core radius (d_smax as written
// by launch_minmax_reduce: sigma = d_smax[0] = d_smax[1]).
int launch_smooth_coef(const NodeRec *d_nodes, int node0, int node1, const int32_t *d_nfirst,
                       const double *sx, const double *sy, const double *sg,
                       const double *d_smax, double eta, int smin, float4 *d_out, hipStream_t s);
// Smooth runs: for every internal node above the smooth level (region >
// `region`) that holds a smooth cell (smooth level, >= smin blobs): the index
// range [x, y) of its smooth-level descendants (consecutive: each level is
// stored in Morton order) if all its blobs lie in them, else (-2, -2);
// (-1, -1) for every other node.  d_flag: 2 nnodes bytes of scratch.  The
// first nroot_slots nodes are roots.
int launch_smooth_ranges(const NodeRec *d_nodes, const int32_t *d_gparent, int nnodes,
                         int nroot_slots, double region, int smin, uint8_t *d_flag, int2 *d_range,
                         hipStream_t s);

// ---- evaluation (eval_kernels.hip) ---------------------------------------
struct TreeView {
  const NodeRec *nodes;
  const int32_t *gparent;
  const double2 *coef;
  const double *sx, *sy, *sg, *ss;
  int nterms;
  double theta;
  // forests: the root of sorted blob t is key[t] >> root_shift (< 0: single tree)
  const uint64_t *key;
  int root_shift;
};
// Treecode velocities at targets (txy interleaved; or the sorted blobs
// themselves when txy is null, then results are scattered through perm).
// Treecode velocities at all n blobs, one wave per 64 consecutive blobs in
// leaf order walking the union of its targets' opened nodes (bitwise those of
// launch_parity_eval at the blobs).
int launch_parity_leaves(const TreeView &tv, int n, const uint32_t *d_perm, double *d_uv,
                         hipStream_t s);
int launch_parity_eval(const TreeView &tv, int nt, const double *d_txy, const uint32_t *d_perm,
                       double *d_uv, hipStream_t s);
// Interaction lists of targets: node index (>= 0) and kind (0 P2P, 1 M2S).
int launch_parity_lists(const TreeView &tv, int nt, const double *d_txy, int cap, int32_t *d_list,
                        int32_t *d_len, hipStream_t s);
// Per target: (far cells with blobs, near pairs); txy null = the blobs themselves.
int launch_parity_counts(const TreeView &tv, int nt, const double *d_txy, int32_t *d_counts,
                         hipStream_t s);
// FP64 direct sum over all sources (style 0: biot_b rule, 1: direct_evaluate rule).
// Sources are split into direct_chunks(ns, nt) chunks (grid y); d_ws holds
// chunks * nt partial sums, reduced in chunk order.
int direct_chunks(int ns, int nt);
int launch_direct(int ns, const double *d_sxy, const double *d_cir, const double *d_cor, int nt,
                  const double *d_txy, double *d_uv, int style, double2 *d_ws, size_t ws_len,
                  hipStream_t s);

// ---- accurate fast mode (fmm_kernels.hip) ------------------------------
// One chunk of the near work of a target leaf with much of it: the tiles
// from position cum up to end, a position being (ordinal of the near item in
// walk order) << 32 | tile; the continuation pass walks the leaf again and
// sums them.
struct SplitCk {
  int64_t cum, end;
  int32_t li, pad;  // target leaf (-1: skip)
};

struct FmmArgs {
  const NodeRec *nodes;
  const int32_t *gparent;
  const double2 *coef;  // unscaled a_k, p per node
  const double *sx, *sy, *sg;
  const float4 *loc;    // (x - c_leaf, y - c_leaf, cir w, w) in FP32, w = log2(e)/sigma^2
  const double4 *loc64; // (x, y, cir, 1/sigma^2) in FP64
  const int32_t *leaves;
  int nleaves;
  // launch order of the target leaves (null: leaf-list order); a fused-pass
  // launch takes positions [pos0, pos0 + npos) of it
  const uint32_t *lorder;
  int pos0, npos;
  const uint32_t *perm;
  double2 *uv;  // caller order
  int p;
  double theta, rcut;
  // forests (many independent systems): root = key[first] >> root_shift;
  // root_shift < 0 for a single tree (root 0)
  const uint64_t *key;
  int root_shift;
  // per target leaf (index in `leaves`): (terms, near radius in core radii);
  // null = p and rcut for every leaf.  smax = largest core radius.
  const float2 *leaf_par;
  double smax;
  // largest w = log2(e) / sigma^2 of the sources (mixed form); with one core
  // radius sw = sqrt(w) scales the FP32 coordinates (0: radii differ)
  float wmax;
  double sw;
  // first blob (in leaf order) of every node's subtree
  const int32_t *nfirst;
  // splitting of target leaves with much near work (split_budget 0: off):
  // checkpoint records (ck_cap of them; *ck_count taken), per target leaf
  // (first record, records), 64 partial sums per record
  int split_budget;
  int ck_cap;
  int ck_base;  // index of record ck[0] among all records (part, leaf_ck)
  SplitCk *ck;
  uint32_t *ck_count;
  int2 *leaf_ck;
  double2 *part;
  // the split leaves (those with records) and their number
  int32_t *split_list;
  uint32_t *split_count;
  // smooth cells (smooth_region 0: off): a near node of region in
  // (smooth_region/2, smooth_region] with at least smooth_min blobs is summed
  // through its coefficients scoef[node * kSmoothF4 ...]
  double smooth_region;
  int smooth_min;
  const float4 *scoef;
  // per node: smooth-level descendants [x, y) of a node above that level
  // whose blobs all lie in them; x = -2: holds smooth cells and other blobs;
  // x = -1: no smooth cell below (see launch_smooth_ranges)
  const int2 *srange;
  // prefix sums over the nodes of the work units of smooth-level nodes (a
  // smooth cell kSmoothUnits, a small leaf its blobs + kStepUnits; nnodes + 1)
  const uint32_t *sunits;
  // far cells listed before a burst sums them (64..kFarList; a walk step
  // adds at most 64)
  int far_burst;
  // local expansions inherited from the parent (l2l = the fewest target-leaf
  // children of a parent that takes part, 2..4; 0: off): per sibling group
  // g, the parent's scaled local expansion plocal[16 g ...]; nodeli[node] =
  // index in `leaves` of a target leaf; *ngroups = sibling groups (nnodes / 4)
  int l2l;
  double2 *plocal;
  const int32_t *nodeli;
  const uint32_t *ngroups;
  int ngroups_h;  // the same count on the host (sizes a grid; 0: read it on the device only)
  // where the walks of a parent's target-leaf children start (written by the
  // parent's walk): per sibling group g, pfcnt[g] entries pfront[kFrontCap g
  // ...], each four siblings (first-child index | skipped-children mask << 27);
  // pfcnt[g] < 0: no list (no parent expansion, or more than front_cap <=
  // kFrontCap entries), the children walk from the root.  pfront null: off.
  uint32_t *pfront;
  int32_t *pfcnt;
  int front_cap;
};
constexpr int kFrontCap = 128;
constexpr int kSkipBits = 27;  // node index bits of a start-list entry
// Node -> leaf-list index of the target leaves, and the group count on the
// device (the inputs of the parent expansions).
int launch_l2l_setup(const int32_t *d_leaves, int nleaves, int32_t *d_nodeli,
                     uint32_t *d_ngroups, int ngroups, hipStream_t s);
// Parent local expansions of the L2L form (no-op with l2l off); d_pcount
// (optional, 3 per sibling group) gets each group's far cells, the parent
// expansion's terms and its target-leaf children.  nblocks x 256 threads.
int launch_parent_local(const FmmArgs &args, bool mixed, int64_t *d_pcount, int nblocks,
                        hipStream_t s);
// Work units of a smooth cell, and of a step's fixed cost (fmm_kernels.hip).
constexpr int kSmoothUnits = 60;
constexpr int kStepUnits = 8;
// Per-node work units of the smooth level (region = 2 eta d_smax[0]) summed
// into d_out[0..nnodes] (exclusive prefix, total last); d_tmp: nnodes u32 of
// scratch.
int launch_smooth_units(const NodeRec *d_nodes, int nnodes, const double *d_smax, double eta,
                        int smin, uint32_t *d_tmp, uint32_t *d_out, hipStream_t s);
int launch_mfma_selftest(const double *a, const double *b, double *out, double *ref, int fp32,
                         hipStream_t s);
// Leaf-relative sources of the *d_count (at most bound) target leaves.  sw =
// d_smax[3] > 0 (all core radii equal, sw = sqrt(log2(e) / sigma^2)): FP32
// sources as (x', y', cir sw, 1) with coordinates relative to the leaf centre
// times sw.
int launch_leaf_local(const NodeRec *d_nodes, const int32_t *d_leaves, const uint32_t *d_count,
                      int bound, const double *sx, const double *sy, const double *sg,
                      const double *ss, float4 *loc, double4 *loc64, const double *d_smax,
                      hipStream_t s);
// Non-empty leaves in blob (Morton) order; their number lands in *d_count.
int launch_leaf_list(const NodeRec *d_nodes, int nnodes, int n, int32_t *d_at, uint32_t *d_flag,
                     uint32_t *d_pos, int32_t *d_leaves, uint32_t *d_count, hipStream_t s);
// Sort keys of the launch order of the target leaves (front leaves: a half
// diagonal above the near radius cut * d_smax[0], the largest first; all
// others after them in leaf-list order), d_val[li] = li, for li < *d_count
// (at most n_bound); sorted by their low 8 bits, stably, they give the order
// in which the evaluation takes the leaves.  *d_front = the front leaves.
int launch_leaf_launch_keys(const NodeRec *d_nodes, const int32_t *d_leaves, int n_bound,
                            const uint32_t *d_count, const double *d_smax, double cut,
                            uint64_t *d_key, uint32_t *d_val, uint32_t *d_front, hipStream_t s);
// d_out[0] = max, d_out[1] = min of n values >= 0 (core radii); d_out[2] =
// w_max = log2(e) / min^2 (capped at 7.2e19), d_out[3] = sqrt(w_max) if max ==
// min, else 0.  Four doubles.
int launch_minmax_reduce(const double *d_v, int n, double *d_out, hipStream_t s);
// First blob of every node's subtree: the first blob of its leftmost leaf.
int launch_node_first_blob(const NodeRec *d_nodes, int nnodes, int32_t *d_out, hipStream_t s);
// One wave per target leaf; with splitting on, d_split_ctr holds five
// counters (records taken and records summed, of the two record regions, and
// the split leaves listed in args.split_list) and
// the continuation passes run on split_blocks x 256 threads.  With a side
// stream and 0 < nfront < nleaves, the first nfront leaves of the launch
// order and their continuation pass run on s (record region 0) while the
// other leaves run on `side` (region 1, records from args.ck_cap on); s then
// waits for side (events fork and join), sums region 1's records and adds the
// chunk sums.  mid (optional): an event recorded before the last continuation
// pass.  ready1, ready (optional): events s waits for before the parent
// expansions and before the fused pass (the build's work on the side stream).
// side_blocks: the grid (x 256 threads) of region 1's continuation pass.
int launch_fmm_eval(const FmmArgs &args, bool mixed, const uint32_t *d_nleaves,
                    uint32_t *d_split_ctr, int split_blocks, hipStream_t s,
                    hipEvent_t mid = nullptr, hipStream_t side = nullptr,
                    hipEvent_t fork = nullptr, hipEvent_t join = nullptr, int nfront = 0,
                    hipEvent_t ready = nullptr, int side_blocks = 0, hipEvent_t ready1 = nullptr);

// ---- time stepping around the evaluation (step_kernels.hip) ---------------
int launch_drift_ab3(int n, double *d_xy, const double *d_hist, double dt, int mirror, double dm,
                     hipStream_t s);
int launch_history_shift(int n, double *d_hist, int first, hipStream_t s);
int launch_store_velocity(int n, const double *d_xy, const double *d_uv, double *d_hist,
                          double uinf, double vinf, double omz, hipStream_t s);
int launch_core_spread(int n, double *d_cor, double dt, double re, hipStream_t s);
int launch_mirror_images(int n, const double *d_xy, const double *d_cir, double dm, double *d_oxy,
                         double *d_ocir, double *d_ocor, hipStream_t s);

// ---- learned per-leaf parameters (learned_kernels.hip) --------------------
constexpr int kMlpWeights = 4866;  // 8 -> 64 -> 64 -> 2
// d_cir: the n circulations in leaf order; d_leaf_abs: per node, the sum of
// |cir| over a leaf's blobs (launch_s2m); d_mean_abs: kMeanAbsWords doubles,
// [0] gets mean |cir| (computed in a fixed order), the rest is scratch for
// block partials.
constexpr int kMeanAbsWords = 1 + 1024;
int launch_leaf_features(const FmmArgs &a, int leaf_cap, const double *d_cir, int n,
                         const double *d_leaf_abs, double *d_mean_abs, float *d_feat,
                         hipStream_t s);
int launch_mlp_forward(const float *d_w, int n, const float *d_x, float *d_out, hipStream_t s);
size_t mlp_scratch_floats(int n);
int launch_mlp_grad(const float *d_w, int n, const float *d_x, const float *d_y, float *d_scratch,
                    float *d_grad, float *d_loss, hipStream_t s);
int launch_adam(float *d_w, const float *d_g, float *d_m, float *d_v, float lr, int t,
                hipStream_t s);
int launch_leaf_params(const float *d_y, int n, float p_margin, float cut_margin, float cut_min,
                       int p_max, float cut_fallback, float2 *d_par, hipStream_t s);
int launch_blob_leaf(const FmmArgs &a, int32_t *d_out, hipStream_t s);
// Persistent grid of nblocks x 256 threads over a work counter; the number of
// target leaves is read from *d_nleaves (graph-capturable).
int launch_fmm_eval_persistent(const FmmArgs &args, bool mixed, const uint32_t *d_nleaves,
                               uint32_t *d_counter, int nblocks, uint32_t *d_split_ctr,
                               int split_blocks, hipStream_t s);
// Per target leaf kCountWords counts (far cells, pairs, smooth evaluations,
// expansion terms, targets).
constexpr int kCountWords = 5;
int launch_fmm_count(const FmmArgs &args, int64_t *d_out, hipStream_t s);
// Accurate modes at nt arbitrary points (interleaved xy): lane per point,
// far test with the point as a target leaf of size 0, FP64 (single tree).
int launch_point_eval(const FmmArgs &args, int nt, const double *d_txy, double *d_uv,
                      hipStream_t s);

}  // namespace fmm2d
