/*
 * oracle2d.h -- CPU oracle for the 2-D vortex-blob treecode (TEST INFRASTRUCTURE).
 *
 * This is the checker, not the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It restates the reference
 * treecode of ver-c/src (quadtree.c, fmm.c, morton_order.c) in plain C with
 * sparse node storage, plus the O(N^2) direct sum of ver-c/src/includes/d_biot.f90.
 */
#ifndef ORACLE2D_H
#define ORACLE2D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tree orc_tree;

/* Parameter file reader; same fscanf formats as quadtree.c:12-27.
 * Returns the number of fields read (0..4), -1 if the file cannot be opened. */
int orc_read_params(const char *path, int *nterms, int *num_p, double *theta, double *r_min);

/* Build the tree and run S2M + M2M (make_tree, quadtree.c:102-171).
 * xy is interleaved x0,y0,x1,y1,...  max_depth caps the level of any node.
 * Returns NULL on error; orc_last_error() then says why. */
orc_tree *orc_build(int n, const double *xy, const double *cir, const double *cor,
                    int nterms, int num_p, double theta, int max_depth);
void orc_free(orc_tree *t);
const char *orc_last_error(void);

/* Tree introspection: nodes in ascending global id (level-major, Morton order). */
int orc_num_nodes(const orc_tree *t);
int orc_depth(const orc_tree *t);
void orc_nodes(const orc_tree *t, uint64_t *gid, int32_t *level, double *cx, double *cy,
               double *region, int32_t *count, int32_t *is_leaf);
/* Blob ids of every node in node order: offsets[nnodes+1] into ids[]; only the
 * ranges of leaves are a partition of 0..n-1 (internal ranges overlap them). */
void orc_node_members(const orc_tree *t, int32_t *lo, int32_t *ids);
/* Multipole coefficients: nnodes x nterms complex, as interleaved (re, im). */
void orc_coeffs(const orc_tree *t, double *re_im);

/* Treecode velocity at targets (biot_tree, fmm.c:69-118); nthreads >= 1. */
void orc_eval(const orc_tree *t, int nt, const double *txy, double *uv, int nthreads);
/* Interaction list of one target in visit order: kind 0 = P2P leaf, 1 = M2S cell.
 * Returns the list length (entries beyond cap are counted but not stored). */
int orc_list(const orc_tree *t, double x, double y, uint64_t *gid, int32_t *kind, int cap);

/* FP64 direct sum over all sources (d_biot.f90:2-64 biot_b without mirror when
 * style == 0: cut-off 1e-6, scale 0.5/pi at the end; style == 1: the
 * direct_evaluate rule of fmm.c:121-142 over all sources, cut-off 1e-10). */
void orc_direct(int ns, const double *xy, const double *cir, const double *cor,
                int nt, const double *txy, double *uv, int style, int nthreads);

/* Morton helpers restated from morton_order.c:4-51 (64-bit ids). */
int64_t orc_global2local(int64_t g, int level);
int64_t orc_local2global(int64_t m, int level);
int orc_get_level(int64_t g);
void orc_get_child(int64_t g, int64_t c[4]);
int64_t orc_get_parent(int64_t g);
int orc_quadrant_pos(const double pos[2], const double cpos[2]);

#ifdef __cplusplus
}
#endif
#endif
