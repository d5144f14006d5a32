/*
 * oracle2d.c -- CPU oracle for the 2-D vortex-blob treecode.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * 2dfmm_amd/: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it; the product never does.
 *
 * It restates the reference treecode of ver-c/src with sparse node storage
 * (the reference allocates a dense complete quadtree, quadtree.c:136-143):
 *   - parameters ............ quadtree.c:12-27   (orc_read_params)
 *   - root box .............. quadtree.c:81-99   (root_box)
 *   - quadrant / centres .... quadtree.c:29-35, 61-79
 *   - top-down stable split . quadtree.c:173-251 (split_all)
 *   - global/local ids ...... morton_order.c:4-51
 *   - S2M ................... fmm.c:18-33
 *   - binomial .............. fmm.c:8-15 (as a table of the same integers)
 *   - M2M ................... fmm.c:35-58
 *   - M2S ................... fmm.c:60-67
 *   - traversal ............. fmm.c:69-118
 *   - P2P ................... fmm.c:121-142
 *   - direct sum ............ ver-c/src/includes/d_biot.f90:2-64 (biot_b)
 * The floating-point expressions are written in the same form as the
 * reference so that, built by the same compiler with the same flags (gcc -O2,
 * no FMA contraction), every coefficient and velocity is bit-identical; the
 * tests check that against the reference build in oracle/_ref.
 */
#include "oracle2d.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint64_t gid;     /* m + (4^l - 1)/3, morton_order.c:10-14 */
  uint64_t local;   /* m: 4*m_parent + q */
  int level;
  double c[2];      /* centre */
  double region;    /* side length */
  int lo, cnt;      /* range in perm[] (stable: ascending blob id) */
  int is_leaf;
  int child;        /* index of child q=0 in nodes[], -1 if none */
} onode;

struct orc_tree {
  int n, nterms, num_p, depth;
  double theta;
  double *pos;      /* [n][2] */
  double *cir, *cor;
  int nnodes, cap;
  onode *nodes;     /* BFS order == ascending gid */
  int *perm;        /* members of every node: leaves partition 0..n-1 */
  double complex *coef;
};

static char g_err[256];
const char *orc_last_error(void) { return g_err; }

/* ---- morton_order.c:4-51, restated on 64-bit ids ---------------------- */
static uint64_t level_offset(int level) { /* (4^l - 1)/3 */
  return ((((uint64_t)1) << (2 * level)) - 1) / 3;
}
int64_t orc_global2local(int64_t g, int level) { return g - (int64_t)level_offset(level); }
int64_t orc_local2global(int64_t m, int level) { return m + (int64_t)level_offset(level); }
int orc_get_level(int64_t g) {
  /* getLevel (morton_order.c:16-29): subtract 4^l until negative */
  int level = 0;
  int64_t value = g;
  for (;;) {
    value -= ((int64_t)1) << (2 * level);
    if (value >= 0) level++; else break;
  }
  return level;
}
void orc_get_child(int64_t g, int64_t c[4]) {
  int level = orc_get_level(g);
  int64_t m = orc_global2local(g, level) << 2;
  for (int i = 0; i < 4; i++) c[i] = orc_local2global(m + i, level + 1);
}
int64_t orc_get_parent(int64_t g) {
  int level = orc_get_level(g);
  int64_t m = orc_global2local(g, level) >> 2;
  return orc_local2global(m, level - 1);
}
int orc_quadrant_pos(const double pos[2], const double cpos[2]) {
  /* quadtree.c:29-35: x >= cx adds 1, y <= cy adds 2 (y measured downward) */
  int q = 0;
  if (pos[0] >= cpos[0]) q++;
  if (pos[1] <= cpos[1]) q += 2;
  return q;
}

/* ---- quadtree.c:12-27 --------------------------------------------------- */
int orc_read_params(const char *path, int *nterms, int *num_p, double *theta, double *r_min) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  int got = 0;
  if (fscanf(f, "%i    nterms", nterms) == 1) got++;
  if (fscanf(f, "%i   num_p", num_p) == 1) got++;
  if (fscanf(f, "%lf theta", theta) == 1) got++;
  if (fscanf(f, "%lf  r_min ! for Tree", r_min) == 1) got++;
  fclose(f);
  return got;
}

/* ---- quadtree.c:37-59, 81-99: root square ------------------------------- */
static void root_box(int n, const double *pos, double c[2], double *region) {
  double mx[2], mn[2];
  for (int d = 0; d < 2; d++) {
    mx[d] = pos[d];
    mn[d] = pos[d];
    for (int i = 1; i < n; i++) {
      double a = pos[2 * i + d];
      if (a > mx[d]) mx[d] = a;
      if (a < mn[d]) mn[d] = a;
    }
  }
  double rx = fabs(mx[0] - mn[0]) + 0.0001;
  double ry = fabs(mx[1] - mn[1]) + 0.0001;
  *region = rx > ry ? rx : ry;
  c[0] = 0.5 * (mx[0] + mn[0]);
  c[1] = 0.5 * (mx[1] + mn[1]);
}

/* quadtree.c:61-79: child centre from parent centre and parent region */
static void child_centre(int q, const double pc[2], double region, double cc[2]) {
  if (q % 2 != 0) cc[0] = pc[0] + 0.25 * region;
  else            cc[0] = pc[0] - 0.25 * region;
  if (q / 2 == 0) cc[1] = pc[1] + 0.25 * region;
  else            cc[1] = pc[1] - 0.25 * region;
}

static int push_node(orc_tree *t) {
  if (t->nnodes == t->cap) {
    int nc = t->cap ? 2 * t->cap : 64;
    onode *nn = (onode *)realloc(t->nodes, sizeof(onode) * (size_t)nc);
    if (!nn) return -1;
    t->nodes = nn;
    t->cap = nc;
  }
  return t->nnodes++;
}

/* quadtree.c:173-251: split the root unconditionally, then every child with
 * num_vor >= num_p; all four children exist; leaf iff count < num_p.
 * Processed breadth first so the node array ends up in ascending gid. */
static int split_all(orc_tree *t, int max_depth) {
  int *tmp = (int *)malloc(sizeof(int) * (size_t)(t->n > 0 ? t->n : 1));
  if (!tmp) return -1;
  for (int s = 0; s < t->nnodes; s++) {
    onode p = t->nodes[s];
    int must_split = (s == 0) || (p.cnt >= t->num_p);
    if (!must_split) continue;
    if (p.level + 1 > max_depth) {
      snprintf(g_err, sizeof g_err,
               "tree depth exceeds %d (>= num_p coincident blobs?)", max_depth);
      free(tmp);
      return -1;
    }
    int cnt[4] = {0, 0, 0, 0};
    for (int i = 0; i < p.cnt; i++) {
      int id = t->perm[p.lo + i];
      cnt[orc_quadrant_pos(&t->pos[2 * id], p.c)]++;
    }
    int off[4];
    off[0] = 0;
    for (int q = 1; q < 4; q++) off[q] = off[q - 1] + cnt[q - 1];
    int fill[4] = {off[0], off[1], off[2], off[3]};
    for (int i = 0; i < p.cnt; i++) {
      int id = t->perm[p.lo + i];
      tmp[fill[orc_quadrant_pos(&t->pos[2 * id], p.c)]++] = id;
    }
    memcpy(&t->perm[p.lo], tmp, sizeof(int) * (size_t)p.cnt);
    int first = -1;
    for (int q = 0; q < 4; q++) {
      int k = push_node(t);
      if (k < 0) { free(tmp); snprintf(g_err, sizeof g_err, "out of memory"); return -1; }
      if (q == 0) first = k;
      onode *c = &t->nodes[k];
      const onode *pp = &t->nodes[s];
      c->level = pp->level + 1;
      c->local = (pp->local << 2) + (uint64_t)q;
      c->gid = c->local + level_offset(c->level);
      child_centre(q, pp->c, pp->region, c->c);
      c->region = 0.5 * pp->region;
      c->lo = pp->lo + off[q];
      c->cnt = cnt[q];
      c->is_leaf = c->cnt < t->num_p;
      c->child = -1;
      if (c->level > t->depth) t->depth = c->level;
    }
    t->nodes[s].child = first;
    t->nodes[s].is_leaf = 0;
  }
  free(tmp);
  return 0;
}

/* fmm.c:8-15 computes C(n,r) by recursion; the same integers from Pascal's
 * triangle, held as doubles (exact for n <= 56). */
static double binom_tab(int n, int k) {
  static double tab[64][64];
  static int init = 0;
  if (!init) {
    for (int a = 0; a < 64; a++) {
      tab[a][0] = 1.0;
      for (int b = 1; b <= a; b++) tab[a][b] = (b == a) ? 1.0 : tab[a - 1][b - 1] + tab[a - 1][b];
    }
    init = 1;
  }
  return tab[n][k];
}

/* fmm.c:18-33: a_k = sum_j cir_j (z_j - z*)^k, k outer, j inner (leaves only) */
static void s2m(orc_tree *t) {
  for (int i = 0; i < t->nnodes; i++) {
    onode *nd = &t->nodes[i];
    if (!nd->is_leaf) continue;
    double complex *a = &t->coef[(size_t)i * t->nterms];
    double complex z_star = nd->c[0] + nd->c[1] * I;
    for (int k = 0; k < t->nterms; k++) {
      for (int j = 0; j < nd->cnt; j++) {
        int id = t->perm[nd->lo + j];
        double complex z_j = t->pos[2 * id] + t->pos[2 * id + 1] * I;
        a[k] += t->cir[id] * cpow((z_j - z_star), k);
      }
    }
  }
}

/* fmm.c:35-58: bottom-up; each parent gathers its non-empty children in
 * quadrant order 3,2,1,0; level-1 nodes are never shifted into the root. */
static void m2m(orc_tree *t) {
  int p = t->nterms;
  double complex coeffs[64];
  for (int s = t->nnodes - 1; s >= 0; s--) {
    onode *par = &t->nodes[s];
    if (par->child < 0 || par->level == 0) continue;
    double complex p_pos = par->c[0] + par->c[1] * I;
    for (int q = 3; q >= 0; q--) {
      int ci = par->child + q;
      onode *ch = &t->nodes[ci];
      if (ch->cnt < 1) continue;
      double complex c_pos = ch->c[0] + ch->c[1] * I;
      const double complex *a = &t->coef[(size_t)ci * p];
      for (int n = 0; n < p; n++) coeffs[n] = 0.0;
      for (int n = 0; n < p; n++)
        for (int k = 0; k < n + 1; k++)
          coeffs[n] += a[k] * binom_tab(n, k) * cpow(-(p_pos - c_pos), n - k);
      double complex *b = &t->coef[(size_t)s * p];
      for (int n = 0; n < p; n++) b[n] += coeffs[n];
    }
  }
}

orc_tree *orc_build(int n, const double *xy, const double *cir, const double *cor,
                    int nterms, int num_p, double theta, int max_depth) {
  g_err[0] = 0;
  if (n < 1) { snprintf(g_err, sizeof g_err, "need at least one blob"); return NULL; }
  if (nterms < 1 || nterms > 60) { snprintf(g_err, sizeof g_err, "nterms out of range"); return NULL; }
  if (num_p < 1) { snprintf(g_err, sizeof g_err, "num_p must be >= 1"); return NULL; }
  if (max_depth > 31) max_depth = 31;
  orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
  t->n = n; t->nterms = nterms; t->num_p = num_p; t->theta = theta;
  t->pos = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  t->cir = (double *)malloc(sizeof(double) * (size_t)n);
  t->cor = (double *)malloc(sizeof(double) * (size_t)n);
  t->perm = (int *)malloc(sizeof(int) * (size_t)n);
  memcpy(t->pos, xy, sizeof(double) * 2 * (size_t)n);
  memcpy(t->cir, cir, sizeof(double) * (size_t)n);
  memcpy(t->cor, cor, sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) t->perm[i] = i;
  int r = push_node(t);
  onode *root = &t->nodes[r];
  memset(root, 0, sizeof *root);
  root_box(n, t->pos, root->c, &root->region);
  root->lo = 0; root->cnt = n; root->is_leaf = 0; root->child = -1;
  if (split_all(t, max_depth) != 0) { orc_free(t); return NULL; }
  t->coef = (double complex *)calloc((size_t)t->nnodes * nterms, sizeof(double complex));
  s2m(t);
  m2m(t);
  return t;
}

void orc_free(orc_tree *t) {
  if (!t) return;
  free(t->pos); free(t->cir); free(t->cor); free(t->perm); free(t->nodes); free(t->coef);
  free(t);
}

int orc_num_nodes(const orc_tree *t) { return t->nnodes; }
int orc_depth(const orc_tree *t) { return t->depth; }

void orc_nodes(const orc_tree *t, uint64_t *gid, int32_t *level, double *cx, double *cy,
               double *region, int32_t *count, int32_t *is_leaf) {
  for (int i = 0; i < t->nnodes; i++) {
    const onode *nd = &t->nodes[i];
    if (gid) gid[i] = nd->gid;
    if (level) level[i] = nd->level;
    if (cx) cx[i] = nd->c[0];
    if (cy) cy[i] = nd->c[1];
    if (region) region[i] = nd->region;
    if (count) count[i] = nd->cnt;
    if (is_leaf) is_leaf[i] = nd->is_leaf;
  }
}

void orc_node_members(const orc_tree *t, int32_t *lo, int32_t *ids) {
  for (int i = 0; i < t->nnodes; i++) lo[i] = t->nodes[i].lo;
  lo[t->nnodes] = t->n;
  memcpy(ids, t->perm, sizeof(int) * (size_t)t->n);
}

void orc_coeffs(const orc_tree *t, double *re_im) {
  for (size_t i = 0; i < (size_t)t->nnodes * t->nterms; i++) {
    re_im[2 * i] = creal(t->coef[i]);
    re_im[2 * i + 1] = cimag(t->coef[i]);
  }
}

/* fmm.c:60-67 */
static void m2s(const orc_tree *t, double complex z_i, double complex *phi, int node) {
  const onode *nd = &t->nodes[node];
  double complex z_star = nd->c[0] + nd->c[1] * I;
  const double complex *a = &t->coef[(size_t)node * t->nterms];
  for (int k = 0; k < t->nterms - 1; k++) *phi += cpow((z_i - z_star), (-k - 1)) * a[k];
}

/* fmm.c:121-142 */
static void p2p(const orc_tree *t, double x, double y, int node, double u_direct[2]) {
  const onode *nd = &t->nodes[node];
  double ax = 0.0, ay = 0.0;
  for (int i = 0; i < nd->cnt; i++) {
    int id = t->perm[nd->lo + i];
    double rx = x - t->pos[2 * id];
    double ry = y - t->pos[2 * id + 1];
    double r = sqrt(rx * rx + ry * ry);
    if (r > 1.0e-10) {
      double rv2 = 1.0 / (r * r);
      double xai = r / t->cor[id];
      double rv = t->cir[id] * (1.0 - exp(-xai * xai));
      ax -= rv * ry * rv2;
      ay += rv * rx * rv2;
    }
  }
  u_direct[0] += 0.5 * ax / M_PI;
  u_direct[1] += 0.5 * ay / M_PI;
}

/* fmm.c:69-118: LIFO traversal; children tested in quadrant order 0..3,
 * opened ones pushed (so popped 3..0), the others evaluated by M2S. */
static void eval_one(const orc_tree *t, double x, double y, double *u, double *v,
                     uint64_t *lg, int32_t *lk, int cap, int *len) {
  double complex phi = 0.0;
  double u_direct[2] = {0.0, 0.0};
  int stack[4 * 32 + 8];
  int top = 0, nl = 0;
  stack[0] = 0;
  double complex z_i = x + y * I;
  while (top >= 0) {
    int ti = stack[top--];
    const onode *nd = &t->nodes[ti];
    if (ti != 0 && nd->is_leaf) {
      if (lg && nl < cap) { lg[nl] = nd->gid; lk[nl] = 0; }
      nl++;
      p2p(t, x, y, ti, u_direct);
    } else {
      for (int q = 0; q < 4; q++) {
        int ci = nd->child + q;
        const onode *ch = &t->nodes[ci];
        double dx = x - ch->c[0], dy = y - ch->c[1];
        double r = sqrt(dx * dx + dy * dy);
        if (ch->region > t->theta * r) {
          stack[++top] = ci;
        } else {
          if (lg && nl < cap) { lg[nl] = ch->gid; lk[nl] = 1; }
          nl++;
          m2s(t, z_i, &phi, ci);
        }
      }
    }
  }
  phi = 0.5 * (-cimag(phi) + creal(phi) * I) / M_PI;
  *u = -creal(phi) + u_direct[0];
  *v = cimag(phi) + u_direct[1];
  if (len) *len = nl;
}

int orc_list(const orc_tree *t, double x, double y, uint64_t *gid, int32_t *kind, int cap) {
  double u, v;
  int len = 0;
  eval_one(t, x, y, &u, &v, gid, kind, cap, &len);
  return len;
}

typedef struct {
  const orc_tree *t;
  const double *txy;
  double *uv;
  int lo, hi;
  /* direct-sum job */
  int ns, style;
  const double *sxy, *cir, *cor;
} job_t;

static void *eval_worker(void *arg) {
  job_t *j = (job_t *)arg;
  for (int i = j->lo; i < j->hi; i++)
    eval_one(j->t, j->txy[2 * i], j->txy[2 * i + 1], &j->uv[2 * i], &j->uv[2 * i + 1], NULL, NULL, 0, NULL);
  return NULL;
}

static void run_jobs(int nt, int nthreads, job_t *proto, void *(*fn)(void *)) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  job_t jobs[256];
  int chunk = (nt + nthreads - 1) / nthreads;
  int used = 0;
  for (int k = 0; k < nthreads; k++) {
    jobs[k] = *proto;
    jobs[k].lo = k * chunk;
    jobs[k].hi = (k + 1) * chunk < nt ? (k + 1) * chunk : nt;
    if (jobs[k].lo >= jobs[k].hi) break;
    used++;
  }
  if (used <= 1) {
    if (used == 1) fn(&jobs[0]);
    return;
  }
  for (int k = 0; k < used; k++) pthread_create(&th[k], NULL, fn, &jobs[k]);
  for (int k = 0; k < used; k++) pthread_join(th[k], NULL);
}

void orc_eval(const orc_tree *t, int nt, const double *txy, double *uv, int nthreads) {
  job_t proto;
  memset(&proto, 0, sizeof proto);
  proto.t = t; proto.txy = txy; proto.uv = uv;
  run_jobs(nt, nthreads, &proto, eval_worker);
}

/* d_biot.f90:21-60 (mirror off) when style == 0; fmm.c:121-142 over all
 * sources when style == 1. */
static void *direct_worker(void *arg) {
  job_t *j = (job_t *)arg;
  const double pitwo = 0.50 / 3.141592653589793238;
  for (int i = j->lo; i < j->hi; i++) {
    double xi = j->txy[2 * i], yi = j->txy[2 * i + 1];
    double ax = 0.0, ay = 0.0;
    double cut = j->style == 0 ? 1.0e-6 : 1.0e-10;
    for (int s = 0; s < j->ns; s++) {
      double rx = xi - j->sxy[2 * s];
      double ry = yi - j->sxy[2 * s + 1];
      double r = sqrt(rx * rx + ry * ry);
      if (r > cut) {
        double rv2 = 1.0 / (r * r);
        double xai = r / j->cor[s];
        double rv = j->cir[s] * (1.0 - exp(-xai * xai));
        if (j->style == 0) {
          ax = ax - ry * rv * rv2;
          ay = ay + rx * rv * rv2;
        } else {
          ax -= rv * ry * rv2;
          ay += rv * rx * rv2;
        }
      }
    }
    if (j->style == 0) {
      j->uv[2 * i] = ax * pitwo;
      j->uv[2 * i + 1] = ay * pitwo;
    } else {
      j->uv[2 * i] = 0.5 * ax / M_PI;
      j->uv[2 * i + 1] = 0.5 * ay / M_PI;
    }
  }
  return NULL;
}

void orc_direct(int ns, const double *xy, const double *cir, const double *cor,
                int nt, const double *txy, double *uv, int style, int nthreads) {
  job_t proto;
  memset(&proto, 0, sizeof proto);
  proto.ns = ns; proto.sxy = xy; proto.cir = cir; proto.cor = cor;
  proto.txy = txy; proto.uv = uv; proto.style = style;
  run_jobs(nt, nthreads, &proto, direct_worker);
}
