/*
 * ref_harness.c -- drives the reference treecode (ver-c/src/*.c compiled in
 * place from /root/reference by oracle/Makefile) so that its tree,
 * coefficients, interaction lists and velocities can be compared with the
 * oracle.  TEST INFRASTRUCTURE ONLY; the reference sources are never copied.
 *
 * The reference reads "../cal_cond/fmm_param.dat" relative to the working
 * directory (quadtree.c:14) and keeps its state in the globals qtree/vors
 * (quadtree.h:49-50); this harness writes the parameter file into a private
 * temporary directory, changes into it for the duration of make_tree, and runs
 * make_tree on a thread with a large stack (calc_region, quadtree.c:83, holds
 * an N-double array on the stack).
 */
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "fmm.h"
#include "morton_order.h"
#include "quadtree.h"

#define EXPORT __attribute__((visibility("default")))

/* fmm.c declares these without a header prototype for M2S. */
void M2S(double complex z_i, double complex *phi, int calculation_node);

static char g_dir[512];
static int g_have_dir = 0;

static int ensure_dir(void) {
  if (g_have_dir) return 0;
  char tmpl[] = "/tmp/ref2dfmm_XXXXXX";
  if (!mkdtemp(tmpl)) return -1;
  snprintf(g_dir, sizeof g_dir, "%s", tmpl);
  char p[600];
  snprintf(p, sizeof p, "%s/cal_cond", g_dir);
  if (mkdir(p, 0700) != 0) return -1;
  snprintf(p, sizeof p, "%s/run", g_dir);
  if (mkdir(p, 0700) != 0) return -1;
  g_have_dir = 1;
  return 0;
}

typedef struct {
  int n;
  double *xy, *cir, *cor;
} mt_args;

static void *mt_thread(void *arg) {
  mt_args *a = (mt_args *)arg;
  make_tree(a->n, (double(*)[2])a->xy, a->cir, a->cor);
  return NULL;
}

/* Runs the reference make_tree with the given parameters (written to the
 * parameter file in the reference's own text format, fmm_param.dat). */
EXPORT int ref_make_tree(int n, const double *xy, const double *cir, const double *cor,
                         int nterms, int num_p, double theta) {
  if (ensure_dir() != 0) return -1;
  char p[600];
  snprintf(p, sizeof p, "%s/cal_cond/fmm_param.dat", g_dir);
  FILE *f = fopen(p, "w");
  if (!f) return -2;
  fprintf(f, "%d    nterms\n%d   num_p\n%.17g theta\n0.01  r_min ! for Tree\n", nterms, num_p, theta);
  fclose(f);
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) return -3;
  snprintf(p, sizeof p, "%s/run", g_dir);
  if (chdir(p) != 0) return -4;
  mt_args a;
  a.n = n;
  a.xy = (double *)xy;
  a.cir = (double *)cir;
  a.cor = (double *)cor;
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, (size_t)64 * 1024 * 1024 + (size_t)n * 16);
  pthread_t th;
  int rc = pthread_create(&th, &attr, mt_thread, &a);
  if (rc == 0) pthread_join(th, NULL);
  pthread_attr_destroy(&attr);
  if (chdir(cwd) != 0) return -5;
  return rc == 0 ? 0 : -6;
}

EXPORT int ref_nterms(void) { return qtree.nterms; }
EXPORT int ref_num_p(void) { return qtree.num_p; }
EXPORT double ref_theta(void) { return qtree.theta; }
EXPORT int ref_dense_nodes(void) { return qtree.num_nodes; }

/* Existing nodes of the dense array: the root, and every child of an
 * existing non-leaf node (split_cell creates all four children). */
static unsigned char *exists_map(void) {
  int nn = qtree.num_nodes;
  unsigned char *ex = (unsigned char *)calloc((size_t)nn, 1);
  ex[0] = 1;
  for (int i = 1; i < nn; i++) {
    int p = getParent(i);
    ex[i] = ex[p] && !qtree.node[p].is_leaf;
  }
  return ex;
}

EXPORT int ref_num_nodes(void) {
  unsigned char *ex = exists_map();
  int c = 0;
  for (int i = 0; i < qtree.num_nodes; i++) c += ex[i];
  free(ex);
  return c;
}

/* Existing nodes in ascending global id. For the root, is_leaf is reported 0. */
EXPORT void ref_nodes(uint64_t *gid, int32_t *level, double *cx, double *cy, double *region,
                      int32_t *count, int32_t *is_leaf, int32_t *lo, int32_t *members,
                      double *coeffs) {
  unsigned char *ex = exists_map();
  int k = 0, off = 0, p = qtree.nterms;
  for (int i = 0; i < qtree.num_nodes; i++) {
    if (!ex[i]) continue;
    Node *nd = &qtree.node[i];
    gid[k] = (uint64_t)i;
    level[k] = i == 0 ? 0 : nd->level;
    cx[k] = nd->center_pos[0];
    cy[k] = nd->center_pos[1];
    region[k] = nd->region;
    count[k] = nd->num_vor;
    is_leaf[k] = i == 0 ? 0 : nd->is_leaf;
    lo[k] = off;
    if (is_leaf[k]) {
      for (int j = 0; j < nd->num_vor; j++) members[off + j] = nd->vor_id[j];
      off += nd->num_vor;
    }
    for (int j = 0; j < p; j++) {
      coeffs[(size_t)k * 2 * p + 2 * j] = creal(nd->outer_coeffs[j]);
      coeffs[(size_t)k * 2 * p + 2 * j + 1] = cimag(nd->outer_coeffs[j]);
    }
    k++;
  }
  lo[k] = off;
  free(ex);
}

/* The reference's own biot_tree for each target. */
EXPORT void ref_eval(int nt, const double *txy, double *uv) {
  for (int i = 0; i < nt; i++) biot_tree(txy[2 * i], txy[2 * i + 1], &uv[2 * i], &uv[2 * i + 1]);
}

/* The decision rule of fmm.c:89-112 walked over the reference's own tree,
 * calling the reference's M2S and direct_evaluate in the order it visits
 * nodes; u, v are then formed as fmm.c:115-117 does.  Equality of (u, v)
 * with ref_eval shows that the recorded list is the one biot_tree follows. */
EXPORT int ref_list(double x, double y, uint64_t *gid, int32_t *kind, int cap, double *u, double *v) {
  double complex phi = 0.0;
  double u_direct[2] = {0.0, 0.0};
  int stack[4 * 64];
  int top = 0, nl = 0;
  stack[0] = 0;
  double complex z_i = x + y * I;
  while (top >= 0) {
    int t = stack[top--];
    if (qtree.node[t].is_leaf) {
      if (nl < cap) { gid[nl] = (uint64_t)t; kind[nl] = 0; }
      nl++;
      direct_evaluate(x, y, t, u_direct);
    } else {
      int c[4];
      getChild(t, c);
      for (int q = 0; q < 4; q++) {
        double dx = x - qtree.node[c[q]].center_pos[0];
        double dy = y - qtree.node[c[q]].center_pos[1];
        double r = sqrt(dx * dx + dy * dy);
        if (qtree.node[c[q]].region > qtree.theta * r) {
          stack[++top] = c[q];
        } else {
          if (nl < cap) { gid[nl] = (uint64_t)c[q]; kind[nl] = 1; }
          nl++;
          M2S(z_i, &phi, c[q]);
        }
      }
    }
  }
  phi = 0.5 * (-cimag(phi) + creal(phi) * I) / M_PI;
  *u = -creal(phi) + u_direct[0];
  *v = cimag(phi) + u_direct[1];
  return nl;
}

/* Morton helpers of morton_order.c, exposed for known-answer tests. */
EXPORT int ref_getLevel(int g) { return getLevel(g); }
EXPORT int ref_getParent(int g) { return getParent(g); }
EXPORT void ref_getChild(int g, int c[4]) { getChild(g, c); }
EXPORT int ref_global2local(int g, int l) { return global2local(g, l); }
EXPORT int ref_local2global(int m, int l) { return local2global(m, l); }
EXPORT int ref_quadrant_pos(double x, double y, double cx, double cy) {
  double a[2] = {x, y}, b[2] = {cx, cy};
  return quadrant_pos(a, b);
}
