/*
 * fmm2d.h -- C ABI of the MI355X (gfx950) 2-D vortex-blob tree library.
 *
 * Drop-in for the reference treecode's two entry points, which the vortex
 * solver binds through ver-c/src/m_tree_bindc.f90 (bind(c), no name mangling):
 *
 *   make_tree  replaces ver-c/src/quadtree.c:102-171 (declared quadtree.h:38-41,
 *              bound at m_tree_bindc.f90:4-8)
 *   biot_tree  replaces ver-c/src/fmm.c:69-118 (declared fmm.h:7, bound at
 *              m_tree_bindc.f90:10-14)
 *
 * Everything else is an extension with no counterpart in the reference; each
 * says which reference behaviour it exposes or which caller it serves.
 *
 * Threading: make_tree and every fmm_* call that builds or evaluates must not
 * run concurrently with each other (as in the reference, whose state is
 * global).  biot_tree may be called from several threads after make_tree.
 * Errors: functions returning int give 0 on success and -1 on failure, with
 * fmm_last_error() describing it; the two reference entry points return
 * nothing, so they print the error to stderr, make_tree leaves an empty tree
 * and biot_tree returns u = v = 0 (the reference has no error path at all).
 */
#ifndef FMM2D_H
#define FMM2D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- the reference ABI ------------------------------------------------- */

/* Builds the quadtree over nvor_b blobs and its multipole expansions.
 * vor_b is x0,y0,x1,y1,... (Fortran vor_b(2,nvor_b), ver-c/memo:1-2); cir_b
 * and cor_b are the circulations and core radii.  Inputs are read during the
 * call only.  Parameters come from "../cal_cond/fmm_param.dat" relative to
 * the working directory, re-read on every call, as quadtree.c:12-27 does,
 * unless fmm_set_params() fixed them.  nvor_b == 0 gives an empty tree
 * (velocities 0).  The velocities at all blobs are computed on the GPU right
 * away (the solver asks for exactly those, e_velo_mpi.f90:37-48), so that
 * biot_tree at a blob position is a lookup. */
void make_tree(int nvor_b, const double *vor_b, const double *cir_b, const double *cor_b);

/* Velocity (u, v) induced at (x, y) by the blobs of the last make_tree, by
 * the reference's treecode rule (fmm.c:69-118) in the treecode mode, or by the
 * accurate modes (fmm_set_mode).  A blob position returns the velocity
 * computed for all blobs (during make_tree, or on the first call); any other
 * point is evaluated on the GPU one at a time (use fmm_eval_batch for many
 * points).  Trees built by fmm_make_tree_device or the mirror variant are
 * looked up the same way. */
void biot_tree(double x, double y, double *u, double *v);

/* ---- extensions --------------------------------------------------------- */

/* Fix nterms, num_p, theta, r_min (the four fields of fmm_param.dat) for all
 * later make_tree calls; the parameter file is then no longer read.
 * nterms in [1, 32], num_p >= 1, theta > 0. */
int fmm_set_params(int nterms, int num_p, double theta, double r_min);
/* Go back to reading a parameter file on every make_tree; path NULL means the
 * reference's "../cal_cond/fmm_param.dat". */
int fmm_use_param_file(const char *path);
int fmm_get_params(int *nterms, int *num_p, double *theta, double *r_min);

/* Evaluation mode (no reference counterpart beyond mode 0).
 * FMM2D_MODE_TREECODE is the reference's algorithm (per-target opening test,
 * M2S far field, direct near field), parameters from fmm_param.dat.
 * FMM2D_MODE_FMM64 and FMM2D_MODE_MIXED bound the error against the
 * regularised FP64 direct sum: cells are far only when theta-separated from
 * the target leaf AND farther than cut_sigmas * (largest core radius), far
 * cells enter one local expansion per target leaf through a dense complex
 * contraction on MFMA, near leaves are summed directly.  MIXED runs the near
 * pairs in FP32 (coordinates relative to a leaf centre; groups of pairs
 * summed in FP32 and accumulated in FP64) and the contraction on FP32 MFMA;
 * expansions and the final accumulation stay FP64.  At the blobs
 * (make_tree's prefetch, fmm_eval_sources*, biot_tree at a blob position)
 * these modes use the target-leaf kernels; other points (fmm_eval_batch*,
 * biot_tree elsewhere) get the same far test with the point as a target leaf
 * of size zero, multipoles evaluated at the point and FP64 near sums. */
#define FMM2D_MODE_TREECODE 0
#define FMM2D_MODE_FMM64 1
#define FMM2D_MODE_MIXED 2
int fmm_set_mode(int mode);
/* p in [2, 16] expansion terms, leaf in [1, 64] leaf capacity, theta in (0, 1),
 * cut_sigmas >= 0.  Defaults 16, 64, 0.5, 4. */
int fmm_set_fmm_params(int p, int leaf, double theta, double cut_sigmas);
/* Checks the MFMA lane layout the contraction relies on: a 16x8 by 8x16
 * product through the FP64 (fp32 = 0) or FP32 (fp32 = 1) MFMA path against a
 * plain loop on the device; *max_abs_err gets the largest difference. */
int fmm_selftest_mfma(int fp32, double *max_abs_err);

/* Treecode mode at the blobs: 1 (default) one wave per 64 consecutive blobs in
 * leaf order walking the union of its targets' opened nodes, 0 one lane per
 * target walking alone; the results are the same bits (each target meets its
 * own events in the reference's order either way). */
int fmm_set_treecode_walk(int per_leaf);

/* 0: make_tree does not evaluate at the blobs; 1 (default): it does,
 * asynchronously.  Only the timing changes: with prefetch off, biot_tree at a
 * blob position evaluates every blob on its first call in the accurate modes
 * (the same values as the prefetch) and that one point in the treecode mode
 * (the same rule). */
int fmm_set_prefetch(int on);

const char *fmm_last_error(void);

/* make_tree on device-resident inputs (device pointers, same layout); no
 * velocity prefetch.  For callers whose blobs already live in HBM. */
int fmm_make_tree_device(int n, const double *d_xy, const double *d_cir, const double *d_cor);

/* Batched systems (no reference counterpart; for callers that advance many
 * independent blob systems at once).  The nsys systems are concatenated:
 * system s owns blobs [offsets[s], offsets[s+1]) (offsets on the host, nsys+1
 * entries, offsets[0] = 0).  One pass builds one tree per system (a forest:
 * per-system root squares, keys prefixed by the system id, one sort, one
 * level loop); fmm_eval_sources* then evaluates every blob against its own
 * system only.  Results per system equal those of make_tree on that system
 * alone with the same parameters. */
int fmm_make_forest(int nsys, const int32_t *offsets, const double *xy, const double *cir,
                    const double *cor);
int fmm_make_forest_device(int nsys, const int32_t *offsets, const double *d_xy,
                           const double *d_cir, const double *d_cor);

/* Accurate modes: evaluate with a persistent grid of `blocks` x 4 waves (one
 * wave per workgroup) that take target leaves from a work counter (0 = one
 * wave per leaf over a sized grid), optionally replayed from a captured HIP
 * graph (use_graph = 1). */
int fmm_set_persistent(int blocks, int use_graph);
/* Accurate modes: target leaves with more than `budget` units of near work
 * (a unit = one wave step over 64 target-source pairs, plus 2 per near
 * leaf) are split into chunks of at least `budget` units that run on other
 * waves; their partial sums are added in chunk order, so results are
 * deterministic.  0 turns splitting off.  Default 8192. */
int fmm_set_split(int budget);
int fmm_get_split(void);
/* Accurate modes: far cells a target leaf lists before one MFMA burst sums
 * them into its local expansion (64..128, default 128; a knob of the
 * autotuner: fewer cells per burst, smaller bursts at the end of the walk). */
int fmm_set_far_burst(int cells);
/* Accurate modes: local expansions inherited from the parent.  A far cell
 * that is also far from a target leaf's parent, and that every target-leaf
 * child of that parent would take as far, is summed once into the parent's
 * local expansion, which reaches each of those children through the exact
 * shift L2L (no reference counterpart; the reference's ver-fortran variant
 * shifts locals in fmm.f90:142-168).  min_children: parents with at least
 * this many target-leaf children take part (1 or 2: all that can, 3, 4;
 * default 2); 0: every target leaf contracts all of its far cells itself. */
int fmm_set_l2l(int min_children);
/* Where the walks of the parent-expansion children start (no reference
 * counterpart).  cap > 0 (default 128, the most): the parent's walk lists the
 * sibling groups below the nodes all its target-leaf children open that hold
 * a child neither covered by its expansion nor opened by all of them, and
 * each child's walk starts there; a parent with more than cap such groups,
 * and every parent with cap = 0, leaves its children to walk from the root,
 * skipping the covered cells on the way.  Both meet the same far cells and
 * near items (in another order, so the sums agree to rounding, not
 * bitwise); the lists take fewer steps. */
int fmm_set_start_lists(int cap);
/* Smooth cells (no reference counterpart; mixed mode with one core radius
 * sigma).  A near node whose half side is at most eta sigma and that holds at
 * least min_blobs blobs is summed for every target through a degree-12 Taylor
 * expansion of the regularised kernel about its centre (coefficients computed
 * at make_tree) instead of pair by pair.  eta in [0, 0.6]; 0 turns it off.
 * Takes effect at the next make_tree.  Default eta 0.5, min_blobs 48. */
int fmm_set_smooth(double eta, int min_blobs);
int fmm_get_smooth(double *eta, int *min_blobs);
/* Compute units of the current device (for sizing persistent grids). */
int fmm_device_cus(void);

/* Learned per-leaf parameters (no reference counterpart).  In the accurate
 * modes each target leaf may use its own expansion order p (<= the set p) and
 * near-field radius (in core radii).  With fmm_set_learned(w, ...) every
 * build computes 8 features per target leaf, runs the FP32 MLP 8->64->64->2
 * with weights w (4866 floats: W1[64][8] b1[64] W2[64][64] b2[64] W3[2][64]
 * b3[2], ReLU) and takes p = ceil(11 + 5 y0 + p_margin), cut = 3 + 2 y1 +
 * cut_margin (clamped to [4, p] and [floor, 6], fmm_set_learned_floor); a
 * non-finite output falls back to the fixed p and cut.  w = NULL switches it
 * off. */
int fmm_set_learned(const float *weights, float p_margin, float cut_margin);
/* Floor of the learned near-field radius, in core radii (1..6, default 2.5):
 * a far pair at c core radii carries exp(-c^2) of its own velocity as the
 * neglected core term (0.2% at 2.5), which bounds what a single leaf's
 * prediction can cost. */
int fmm_set_learned_floor(float cut_min);
/* The current tree's target leaves (accurate modes): count; features
 * (count x 8 floats); the target-leaf index of every blob (caller order); the
 * per-leaf (p, cut) in force (count x 2); or set them explicitly (until the
 * next build; NULL clears). */
int fmm_num_target_leaves(void);
int fmm_leaf_features(float *feat);
int fmm_blob_leaf(int32_t *leaf_of_blob);
int fmm_leaf_params(float *par);
int fmm_set_leaf_params(const float *par, int n);
/* The MLP itself, on the GPU: forward pass of n samples (x: n x 8, out: n x 2);
 * gradient of mean((out - y)^2) (grad: 4866 floats, *loss the mean); Adam
 * training on minibatches (weights updated in place, loss_hist[epochs]). */
int fmm_mlp_forward(int n, const float *x, const float *weights, float *out);
int fmm_mlp_grad(int n, const float *x, const float *y, const float *weights, float *grad,
                 float *loss);
int fmm_mlp_train(int n, const float *x, const float *y, float *weights, int epochs, int batch,
                  float lr, unsigned seed, float *loss_hist);

/* Mirror images (biot_b's i_mirror, ver-c/src/includes/d_biot.f90:39-53; the
 * reference treecode has none, README.md:8): the tree is built over the n
 * blobs and their images (x, 2 d_mirror - y) with circulation -cir and the
 * point-vortex kernel (core 0); fmm_eval_sources* then return the n real
 * blobs' velocities.  Device pointers. */
int fmm_make_tree_mirror_device(int n, const double *d_xy, const double *d_cir,
                                const double *d_cor, double d_mirror);

/* Time stepping on device-resident blobs (the solver's updates around the
 * evaluation; no counterpart in the tree code).  hist is laid out like the
 * solver's vor_vb(2,3,n): hist[(i*3 + k)*2 + c], k = 0 newest.
 *   drift: x += dt (23/12 v1 - 4/3 v2 + 5/12 v3) (e_drift.f90:29-44), with
 *          reflection below y = d_mirror when mirror != 0 (e_drift.f90:46);
 *   history shift: v3 <- v2, v2 <- v1; first_step copies v1 to v2 first
 *          (e_velo_update.f90:17-38);
 *   store velocity: v1 <- (uinf - omz y, vinf + omz x) + uv (e_velo_mpi.f90:44-47);
 *   core spreading: sigma <- sqrt(sigma^2 + 2.2418^2 dt / re) (e_core.f90:15-19). */
int fmm_drift_ab3_device(int n, double *d_xy, const double *d_hist, double dt, int mirror,
                         double d_mirror);
int fmm_history_shift_device(int n, double *d_hist, int first_step);
int fmm_store_velocity_device(int n, const double *d_xy, const double *d_uv, double *d_hist,
                              double uinf, double vinf, double omz);
int fmm_core_spread_device(int n, double *d_cor, double dt, double re);

/* Velocities at all blobs of the last tree, in the caller's blob order,
 * interleaved u0,v0,u1,v1,...  Host or device output buffer. */
int fmm_eval_sources(double *uv);
int fmm_eval_sources_device(double *d_uv);

/* Velocities at nt arbitrary points xy (interleaved); host or device buffers. */
int fmm_eval_batch(int nt, const double *xy, double *uv);
int fmm_eval_batch_device(int nt, const double *d_xy, double *d_uv);

/* Tree introspection (for tests and tools).  Nodes are reported in ascending
 * global id g = m + (4^l - 1)/3 of the reference's numbering
 * (morton_order.c:4-14): level-major, Morton order inside a level.  Arrays
 * hold nnodes entries (lo: nnodes + 1; members: nvor_b; coeffs: 2*nnodes*nterms
 * doubles).  is_leaf is 0 for the root.  members lists, for each leaf in node
 * order, its blob ids in ascending order; lo[k] is where node k's list starts
 * (internal nodes have empty lists). */
int fmm_tree_size(int *nnodes, int *depth, int *nleaves);
int fmm_dump_tree(uint64_t *gid, int32_t *level, double *cx, double *cy, double *region,
                  int32_t *count, int32_t *is_leaf, int32_t *lo, int32_t *members, double *coeffs);

/* Interaction lists of nt points, in the order the treecode visits them:
 * gid[t*cap + i] and kind[t*cap + i] (0 = direct sum over a leaf, 1 = M2S of
 * a cell); len[t] is the full length (entries past cap are not stored). */
int fmm_dump_lists(int nt, const double *xy, int cap, uint64_t *gid, int32_t *kind, int32_t *len);

/* FP64 direct sum over all ns sources at nt targets on the GPU: style 0 is
 * biot_b (ver-c/src/includes/d_biot.f90:2-64, cut-off 1e-6, mirror off),
 * style 1 the near-field rule of fmm.c:121-142 (cut-off 1e-10). */
int fmm_direct(int ns, const double *xy, const double *cir, const double *cor, int nt,
               const double *txy, double *uv, int style);
int fmm_direct_device(int ns, const double *d_xy, const double *d_cir, const double *d_cor, int nt,
                      const double *d_txy, double *d_uv, int style);

/* fmm_direct_device without any allocation (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a fixed 16 MiB device array that each GPU holds until
 * the process ends, so nt is limited to 2^20 targets (larger nt: -1 with a
 * message, nothing launched).  Runs on `stream` (a hipStream_t; NULL = the
 * null stream) and returns without waiting.  Results equal
 * fmm_direct_device's bit for bit.  Every call on one GPU shares its array:
 * calls that overlap on one GPU (other streams or threads) race. */
int fmm_direct_device_noalloc(int ns, const double *d_xy, const double *d_cir, const double *d_cor,
                              int nt, const double *d_txy, double *d_uv, int style, void *stream);

/* Measurement (no reference counterpart: the reference has no timers).
 * With timing on, HIP events on the library's stream bracket the phases of
 * the last build and evaluation; fmm_get_timing fills up to 10 values in ms:
 * [0] bbox + keys, [1] radix sort, [2] tree levels + leaf order + gather,
 * [3] S2M + M2M and the accurate modes' set-up (smooth cells, learned
 * parameters), [4] the last evaluation over all blobs, [5] tree levels alone,
 * [6] leaf order + gather, [7] S2M + M2M alone, [8] the fused pass of the
 * accurate modes, [9] their split (continuation) passes; -1 where a phase
 * did not run. */
int fmm_set_timing(int on);
int fmm_get_timing(double *ms, int n);
/* Treecode interactions of the last tree with all blobs as targets: far cells
 * evaluated by M2S (cells holding blobs) and near source-target pairs. */
int fmm_count_interactions(long long *m2s, long long *pairs);
/* Same walk, with the smooth cells of the mixed form counted apart: far
 * cells, pairs summed directly, and (target, smooth cell) evaluations. */
int fmm_count_work(long long *m2s, long long *pairs, long long *smooth);
/* The work of that walk in detail (up to n values): [0] far cells (M2L
 * contractions of target leaves; treecode: M2S evaluations), [1] near pairs,
 * [2] (target, smooth cell) evaluations, [3] sum over target leaves of far
 * cells x the leaf's terms^2 (the contractions' size), [4] sum over target
 * leaves of targets x terms (L2P), [5] target leaves, [6] far cells summed
 * into parent expansions (L2L), [7] their far cells x the parent's terms^2,
 * [8] parent expansions, [9] L2L shifts; accurate modes only for [2..9] (0 in
 * the treecode mode). */
int fmm_count_work_detail(long long *out, int n);
/* The same walk per target leaf (accurate modes): out[5 i + 0..4] = far cells,
 * near pairs, (target, smooth cell) evaluations, terms and targets of target
 * leaf i (nleaves = fmm_num_target_leaves()). */
int fmm_count_work_leaves(long long *out, int nleaves);

/* Wait for all work of the library; release all device memory.  The library
 * works on two non-blocking streams of its own, at the default priority: the
 * evaluation runs its front leaves on one and the other leaves on the other
 * and joins them before it ends; make_tree leaves some set-up on the second,
 * which every later call waits for.  fmm_synchronize waits for both. */
int fmm_synchronize(void);
void fmm_release(void);

#ifdef __cplusplus
}
#endif
#endif
