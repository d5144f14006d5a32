// eval_kernels.hip -- velocity evaluation on the GPU (gfx950).
//
// Treecode (parity) traversal, one lane per target, following fmm.c:69-118:
// a node is expanded by testing its four children in quadrant order 0..3;
// child c is opened iff region_c > theta * |z - z_c| (the same FP64
// operations: two products, a sum, a correctly rounded sqrt, a product), an
// unopened child is evaluated at once by M2S (fmm.c:60-67), and the opened
// children are then visited in quadrant order 3..0, depth first, which is the
// order the reference's LIFO pops them.  Instead of a stack the lane keeps one
// 4-bit mask of not-yet-visited opened children per level (128 bits for 31
// levels, in registers) and climbs through the group parent array, so no
// scratch memory is touched.  Opened leaves are summed directly with the
// regularised blob kernel of fmm.c:121-142.  The complex potential and the
// direct part are combined as fmm.c:115-117.
//
// Direct judge: all sources for every target, the rule of d_biot.f90:21-60
// (style 0) or of fmm.c:121-142 (style 1); LDS tiles of 256 sources.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fmm2d_internal.h"

namespace fmm2d {

struct Slots {
  uint64_t lo = 0, hi = 0;
  __device__ inline uint32_t get(int d) const {
    const int i = d - 1;
    return i < 16 ? (uint32_t)(lo >> (4 * i)) & 15u : (uint32_t)(hi >> (4 * (i - 16))) & 15u;
  }
  __device__ inline void set(int d, uint32_t m) {
    const int i = d - 1;
    if (i < 16) {
      lo = (lo & ~(15ull << (4 * i))) | ((uint64_t)m << (4 * i));
    } else {
      const int s = 4 * (i - 16);
      hi = (hi & ~(15ull << s)) | ((uint64_t)m << s);
    }
  }
};

// 1 / d for d > 0: the hardware reciprocal refined by two Newton steps
// (within an ulp or so of the correctly rounded quotient; the velocities are
// held to 1e-12 of the reference, the opening decisions keep the reference's
// exact arithmetic)
__device__ inline double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = r * fma(-d, r, 2.0);
  return r * fma(-d, r, 2.0);
}

// phi += sum_{k=0}^{p-2} a_k (z - z_c)^{-(k+1)}, Horner in w = 1 / (z - z_c)
// with fused multiply-adds
__device__ inline void m2s_horner(const double2 *__restrict__ a, int p, double dx, double dy,
                                  double &phr, double &phi) {
  if (p < 2) return;
  const double inv = rcp_nr(fma(dx, dx, dy * dy));
  const double wr = dx * inv, wi = -dy * inv;
  double2 t = a[p - 2];
  double sr = t.x, si = t.y;
  for (int k = p - 3; k >= 0; k--) {
    const double2 ak = a[k];
    const double nr = fma(sr, wr, fma(-si, wi, ak.x));
    const double ni = fma(sr, wi, fma(si, wr, ak.y));
    sr = nr;
    si = ni;
  }
  phr += fma(sr, wr, -si * wi);
  phi += fma(sr, wi, si * wr);
}

// One pair of fmm.c:121-142 (regularised blob, cut-off r > 1e-10), source
// (sx, sy, cir, core radius) = s; r from the correctly rounded sqrt as in
// the reference, the divisions by reciprocals; a zero core radius keeps the
// reference's infinite r / 0
__device__ inline void p2p_pair(double x, double y, const double4 &s, double &ax, double &ay) {
  const double rx = x - s.x;
  const double ry = y - s.y;
  const double r = sqrt(fma(rx, rx, ry * ry));
  if (r > 1.0e-10) {
    const double rv2 = rcp_nr(r * r);
    // a zero core radius (mirror images) is a point vortex: r / 0 = inf in
    // the reference, and rcp_nr(0) would turn 0 * inf into NaN
    const double xai = s.w > 0.0 ? r * rcp_nr(s.w) : __builtin_inf();
    const double rv = s.z * (1.0 - exp(-xai * xai));
    ax -= rv * ry * rv2;
    ay += rv * rx * rv2;
  }
}

// fmm.c:121-142 for one leaf
__device__ inline void p2p_leaf(const TreeView &tv, int lo, int cnt, double x, double y,
                                double &ud, double &vd) {
  double ax = 0.0, ay = 0.0;
  for (int j = lo; j < lo + cnt; j++)
    p2p_pair(x, y, make_double4(tv.sx[j], tv.sy[j], tv.sg[j], tv.ss[j]), ax, ay);
  ud += 0.5 * ax / M_PI;
  vd += 0.5 * ay / M_PI;
}

// MODE 0: velocities; 1: interaction lists; 2: interaction counts only
// (far cells with blobs, near source-target pairs) for the flop count.
template <int MODE>
__global__ __launch_bounds__(256) void parity_eval_kernel(TreeView tv, int nt,
                                                          const double2 *__restrict__ txy,
                                                          const uint32_t *__restrict__ perm,
                                                          double2 *__restrict__ uv, int cap,
                                                          int32_t *__restrict__ list,
                                                          int32_t *__restrict__ len) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nt) return;
  double x, y;
  if (txy) {
    const double2 p = txy[t];
    x = p.x;
    y = p.y;
  } else {
    x = tv.sx[t];
    y = tv.sy[t];
  }
  const NodeRec *__restrict__ nodes = tv.nodes;
  const int p = tv.nterms;
  const double theta = tv.theta;
  double phr = 0.0, phi = 0.0, ud = 0.0, vd = 0.0;
  Slots sl;
  int d = 0, cur = 0, nl = 0;
  // root: 0 for a single tree; in a forest, the system of the (sorted) blob
  int node = (tv.root_shift < 0 || txy) ? 0 : (int)(tv.key[t] >> tv.root_shift);
  bool expand = true;
  constexpr bool RECORD = MODE == 1, COUNT = MODE == 2;
  int32_t *my_list = RECORD ? list + (size_t)t * cap * 2 : nullptr;
  int32_t n_m2s = 0, n_pairs = 0;
  for (;;) {
    if (expand) {
      const int base = nodes[node].first;
      uint32_t m = 0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int c = base + q;
        const NodeRec C = nodes[c];
        const double dx = x - C.cx, dy = y - C.cy;
        const double r = sqrt(dx * dx + dy * dy);
        if (C.region > theta * r) {
          m |= 1u << q;
        } else {
          if (RECORD) {
            if (nl < cap) { my_list[2 * nl] = c; my_list[2 * nl + 1] = 1; }
            nl++;
          }
          if (COUNT) {
            n_m2s += node_count(C.cnt) > 0 ? 1 : 0;
          } else if (node_count(C.cnt) > 0) {
            m2s_horner(tv.coef + (size_t)c * p, p, dx, dy, phr, phi);
          }
        }
      }
      d++;
      sl.set(d, m);
      cur = base;
    }
    uint32_t m;
    bool done = false;
    for (;;) {
      m = sl.get(d);
      if (m) break;
      if (d == 1) { done = true; break; }
      cur = tv.gparent[cur >> 2];
      d--;
    }
    if (done) break;
    const int q = 31 - __clz(m);
    sl.set(d, m & ~(1u << q));
    node = (cur & ~3) + q;
    const NodeRec N = nodes[node];
    if (node_internal(N.cnt)) {
      expand = true;
    } else {
      if (RECORD) {
        if (nl < cap) { my_list[2 * nl] = node; my_list[2 * nl + 1] = 0; }
        nl++;
      }
      if (COUNT) n_pairs += node_count(N.cnt);
      else p2p_leaf(tv, N.first, node_count(N.cnt), x, y, ud, vd);
      expand = false;
    }
  }
  // fmm.c:115-117
  const double u = 0.5 * phi / M_PI + ud;
  const double v = 0.5 * phr / M_PI + vd;
  if (RECORD) len[t] = nl;
  if (COUNT) {
    len[2 * t] = n_m2s;
    len[2 * t + 1] = n_pairs;
    return;
  }
  if (uv) {
    const int o = perm ? (int)perm[t] : t;
    uv[o] = make_double2(u, v);
  }
}

// Treecode at the blobs, one wave per 64 consecutive blobs in leaf order (the
// targets: a leaf and its neighbours, which open nearly the same nodes; a
// leaf holds fewer than 64 blobs, so one wave per leaf left half its lanes
// without a target): the wave walks the union of the nodes its targets open,
// depth first in the reference's order -- children tested in quadrant order
// 0..3, opened ones visited 3..0 -- each stack entry carrying the mask of
// lanes that opened it, one root entry per tree among the targets (forests).
// At a popped internal node every lane of the mask tests the four children
// itself and evaluates the ones it does not open by M2S at once; at a
// popped leaf the lanes of the mask sum it directly.  Restricted to one lane
// the events are that target's own walk (parity_eval_kernel) in the same
// order with the same arithmetic, so the velocities are bit for bit the same;
// here the lanes of a leaf walk together, and the four children's
// coefficients (for M2S) and a popped leaf's sources (for the direct sums)
// are staged in LDS, one vector load per lane, and read from there in the
// same order as the per-target walk reads them from memory.  The stack holds
// at most 3 entries per level plus one per tree among the 64 targets.
constexpr int kParStack = 3 * kMaxDepth + 1 + 64;
__global__ __launch_bounds__(64) void parity_leaf_kernel(TreeView tv, int n,
                                                         const uint32_t *__restrict__ perm,
                                                         double2 *__restrict__ uv) {
  __shared__ int s_node[kParStack];
  __shared__ uint64_t s_mask[kParStack];
  __shared__ double2 s_coef[4 * kMaxTerms];  // the four children's a_k
  __shared__ double4 s_src[64];              // (x, y, cir, core radius) of a leaf's blobs
  const int lane = threadIdx.x;
  const NodeRec *__restrict__ nodes = tv.nodes;
  const int p = tv.nterms;
  const double theta = tv.theta;
  {
    const int t = blockIdx.x * 64 + lane;
    const bool act = t < n;
    const double x = act ? tv.sx[t] : 0.0, y = act ? tv.sy[t] : 0.0;
    double phr = 0.0, phi = 0.0, ud = 0.0, vd = 0.0;
    // root: 0 for a single tree; in a forest, the system of the blob; one
    // stack entry per root among the targets (disjoint masks, any order)
    const int root = tv.root_shift < 0 ? 0 : (act ? (int)(tv.key[t] >> tv.root_shift) : -1);
    int ns = 0;
    uint64_t left = __builtin_amdgcn_ballot_w64(act);
    while (left) {
      const int r = __builtin_amdgcn_readlane(root, __builtin_ctzll(left));
      const uint64_t m = __builtin_amdgcn_ballot_w64(act && root == r);
      if (lane == 0) {
        s_node[ns] = r;
        s_mask[ns] = m;
      }
      ns++;
      left &= ~m;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    while (ns > 0) {
      ns--;
      const int node = s_node[ns];
      const uint64_t m = s_mask[ns];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const bool mine = (m >> lane) & 1;
      const NodeRec N = nodes[node];
      if (!node_internal(N.cnt)) {
        // p2p_leaf's arithmetic, sources from LDS in 64-blob chunks
        const int lo = N.first, nc = node_count(N.cnt);
        double ax = 0.0, ay = 0.0;
        for (int j0 = 0; j0 < nc; j0 += 64) {
          if (j0 + lane < nc) {
            const int j = lo + j0 + lane;
            s_src[lane] = make_double4(tv.sx[j], tv.sy[j], tv.sg[j], tv.ss[j]);
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          if (mine) {
            const int jn = min(64, nc - j0);
            for (int j = 0; j < jn; j++) p2p_pair(x, y, s_src[j], ax, ay);
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
        if (mine) {
          ud += 0.5 * ax / M_PI;
          vd += 0.5 * ay / M_PI;
        }
        continue;
      }
      const int base = N.first;
      // the children's coefficients: rows base .. base + 3, consecutive
      for (int e = lane; e < 4 * p; e += 64) s_coef[e] = tv.coef[(size_t)base * p + e];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      uint64_t om[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int c = base + q;
        const NodeRec C = nodes[c];
        bool op = false;
        if (mine) {
          const double dx = x - C.cx, dy = y - C.cy;
          const double r = sqrt(dx * dx + dy * dy);
          if (C.region > theta * r)
            op = true;
          else if (node_count(C.cnt) > 0)
            m2s_horner(s_coef + q * p, p, dx, dy, phr, phi);
        }
        om[q] = __builtin_amdgcn_ballot_w64(op);
      }
      // pushed in quadrant order, so the highest opened quadrant is next
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (om[q]) {
          if (lane == 0) {
            s_node[ns] = base + q;
            s_mask[ns] = om[q];
          }
          ns++;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    // fmm.c:115-117
    if (act) uv[perm[t]] = make_double2(0.5 * phi / M_PI + ud, 0.5 * phr / M_PI + vd);
  }
}

int launch_parity_leaves(const TreeView &tv, int n, const uint32_t *d_perm, double *d_uv,
                         hipStream_t s) {
  if (n <= 0) return 0;
  parity_leaf_kernel<<<(n + 63) / 64, 64, 0, s>>>(tv, n, d_perm, reinterpret_cast<double2 *>(d_uv));
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_parity_eval(const TreeView &tv, int nt, const double *d_txy, const uint32_t *d_perm,
                       double *d_uv, hipStream_t s) {
  if (nt <= 0) return 0;
  parity_eval_kernel<0><<<(nt + 255) / 256, 256, 0, s>>>(
      tv, nt, reinterpret_cast<const double2 *>(d_txy), d_txy ? nullptr : d_perm,
      reinterpret_cast<double2 *>(d_uv), 0, nullptr, nullptr);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_parity_lists(const TreeView &tv, int nt, const double *d_txy, int cap, int32_t *d_list,
                        int32_t *d_len, hipStream_t s) {
  if (nt <= 0) return 0;
  parity_eval_kernel<1><<<(nt + 255) / 256, 256, 0, s>>>(
      tv, nt, reinterpret_cast<const double2 *>(d_txy), nullptr, nullptr, cap, d_list, d_len);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_parity_counts(const TreeView &tv, int nt, const double *d_txy, int32_t *d_counts,
                         hipStream_t s) {
  if (nt <= 0) return 0;
  parity_eval_kernel<2><<<(nt + 255) / 256, 256, 0, s>>>(
      tv, nt, reinterpret_cast<const double2 *>(d_txy), nullptr, nullptr, 0, nullptr, d_counts);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// FP64 direct sum (judge)
// ---------------------------------------------------------------------------
constexpr int kDirTile = 256;

// Block (bx, by): targets bx*256.., sources of chunk by; raw sums per chunk.
__global__ __launch_bounds__(256) void direct_kernel(int ns, int csize,
                                                     const double2 *__restrict__ sxy,
                                                     const double *__restrict__ cir,
                                                     const double *__restrict__ cor, int nt,
                                                     const double2 *__restrict__ txy,
                                                     double2 *__restrict__ part, int style) {
  __shared__ double lx[kDirTile], ly[kDirTile], lg[kDirTile], ls[kDirTile];
  const int t = blockIdx.x * 256 + threadIdx.x;
  double x = 0.0, y = 0.0;
  if (t < nt) {
    const double2 p = txy[t];
    x = p.x;
    y = p.y;
  }
  const double cut = style == 0 ? 1.0e-6 : 1.0e-10;
  const int s0 = blockIdx.y * csize, s1 = min(ns, s0 + csize);
  double ax = 0.0, ay = 0.0;
  for (int b = s0; b < s1; b += kDirTile) {
    __syncthreads();
    const int j = b + threadIdx.x;
    if (j < s1) {
      const double2 q = sxy[j];
      lx[threadIdx.x] = q.x;
      ly[threadIdx.x] = q.y;
      lg[threadIdx.x] = cir[j];
      ls[threadIdx.x] = cor[j];
    }
    __syncthreads();
    const int m = min(kDirTile, s1 - b);
    for (int k = 0; k < m; k++) {
      const double rx = x - lx[k];
      const double ry = y - ly[k];
      const double r = sqrt(rx * rx + ry * ry);
      if (r > cut) {
        const double rv2 = 1.0 / (r * r);
        const double xai = r / ls[k];
        const double rv = lg[k] * (1.0 - exp(-xai * xai));
        if (style == 0) {
          ax = ax - ry * rv * rv2;
          ay = ay + rx * rv * rv2;
        } else {
          ax -= rv * ry * rv2;
          ay += rv * rx * rv2;
        }
      }
    }
  }
  if (t < nt) part[(size_t)blockIdx.y * nt + t] = make_double2(ax, ay);
}

// Chunk sums in chunk order, then the scaling of the chosen rule.
__global__ __launch_bounds__(256) void direct_reduce_kernel(const double2 *__restrict__ part,
                                                            int nchunks, int nt,
                                                            double2 *__restrict__ uv, int style) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nt) return;
  double ax = 0.0, ay = 0.0;
  for (int c = 0; c < nchunks; c++) {
    const double2 v = part[(size_t)c * nt + t];
    ax += v.x;
    ay += v.y;
  }
  if (style == 0) {
    const double pitwo = 0.50 / 3.141592653589793238;
    uv[t] = make_double2(ax * pitwo, ay * pitwo);
  } else {
    uv[t] = make_double2(0.5 * ax / M_PI, 0.5 * ay / M_PI);
  }
}

int direct_chunks(int ns, int nt) {
  const int tb = (nt + 255) / 256;
  int c = (2048 + tb - 1) / tb;
  c = std::min(c, std::max(1, ns / 1024));
  return std::max(c, 1);
}

int launch_direct(int ns, const double *d_sxy, const double *d_cir, const double *d_cor, int nt,
                  const double *d_txy, double *d_uv, int style, double2 *d_ws, size_t ws_len,
                  hipStream_t s) {
  if (nt <= 0) return 0;
  const int nc = direct_chunks(ns, nt);
  if ((size_t)nc * nt > ws_len) {
    set_error("direct sum workspace too small");
    return -1;
  }
  const int csize = ns > 0 ? (ns + nc - 1) / nc : 0;
  direct_kernel<<<dim3((nt + 255) / 256, nc), 256, 0, s>>>(
      ns, csize, reinterpret_cast<const double2 *>(d_sxy), d_cir, d_cor, nt,
      reinterpret_cast<const double2 *>(d_txy), d_ws, style);
  FMM_HIP_TRY(hipGetLastError());
  direct_reduce_kernel<<<(nt + 255) / 256, 256, 0, s>>>(d_ws, nc, nt,
                                                        reinterpret_cast<double2 *>(d_uv), style);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace fmm2d
