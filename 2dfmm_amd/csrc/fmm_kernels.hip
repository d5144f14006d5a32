// fmm_kernels.hip -- accurate fast mode: one wave per target leaf (gfx950).
//
// Unlike the reference treecode (per-target opening test, singular far field
// right next to the target), this mode bounds the error against the
// regularised direct sum:
//   * a source cell C is far from the target leaf A iff
//       R_A + R_C <= theta * d   and   d - R_A - R_C >= r_cut,
//     with R the half diagonal of a cell and d the distance of the centres;
//     r_cut = cut_sigmas * max core radius keeps every far pair beyond the
//     range where the blob core matters (the neglected term is
//     exp(-r^2 / sigma^2) of the pair's velocity);
//   * near leaves are summed directly with the regularised blob kernel;
//   * far cells of A are gathered into ONE dense complex contraction per
//     target leaf on MFMA:
//         G[l][k] = sum_over_far_cells (s_A w)^(l+1) * ((-1)^(k+1) a_k w^k),
//         w = 1 / (c_C - c_A),
//     and the scaled local expansion is  Lt_l = (1/s_A) sum_k C(l+k,k) G[l][k],
//     i.e. the M2L of every far cell of A folded into one rank-K update of a
//     16 x 32 real matrix (K = 2 x far cells); the potential at a target is
//     sum_l Lt_l ((z - c_A)/s_A)^l (L2P).
// The traversal decisions are uniform across the wave (they depend on the
// target leaf only), so the wave walks the tree together; lanes are
// (target, replica) pairs for the direct part.
//
// Precision: FP64 everywhere for the plain form.  The mixed form runs the
// direct part in FP32 on leaf-relative coordinates with Kahan-compensated
// sums (promoted to FP64 per source leaf) and the contraction on FP32 MFMA
// (promoted to FP64 every 8 cells); expansions (S2M/M2M) and the L2P
// accumulation stay in FP64.
#include <hip/hip_runtime.h>

#include "fmm2d_internal.h"

namespace fmm2d {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

constexpr int kFmmMaxP = 16;

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline double2 cpow_u(double2 b, int e) {
  double2 r = make_double2(1.0, 0.0);
  while (e > 0) {
    if (e & 1) r = cmul(r, b);
    e >>= 1;
    if (e) b = cmul(b, b);
  }
  return r;
}

// 1 - exp(-x) for x >= 0 in FP32 without the cancellation of small x: the
// Taylor series to x^8 below 0.5 (truncation < 1e-8 relative), 1 - exp(-x)
// above.  Explicit fused multiply-adds (the build disables contraction).
// The series sits in a branch: a wave skips it whenever none of its lanes
// has a pair closer than 0.7 core radii (most pairs of a near field).
__device__ inline float one_minus_exp_neg(float x) {
  float r = 1.0f - __expf(-x);
  if (x < 0.5f) {
    float q = fmaf(x, -1.0f / 40320.0f, 1.0f / 5040.0f);
    q = fmaf(x, q, -1.0f / 720.0f);
    q = fmaf(x, q, 1.0f / 120.0f);
    q = fmaf(x, q, -1.0f / 24.0f);
    q = fmaf(x, q, 1.0f / 6.0f);
    q = fmaf(x, q, -0.5f);
    q = fmaf(x, q, 1.0f);
    r = x * q;
  }
  return r;
}

__device__ inline void kahan_add(float &s, float &c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// ---------------------------------------------------------------------------
// MFMA helpers: D(16x16) += A(16x4) B(4x16); lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and holds D[row][l%16] in its register r = 0..3, where
//   f64 (v_mfma_f64_16x16x4_f64): row = l/16 + 4 r   (rows interleaved)
//   f32 (v_mfma_f32_16x16x4_f32): row = 4 (l/16) + r (rows blocked)
// (measured with one-hot operands on gfx950; fmm_selftest_mfma checks it).
// ---------------------------------------------------------------------------
__device__ inline d4_t mfma16x16x4(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4_t mfma16x16x4(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline int mfma_row(double, int k, int r) { return k + 4 * r; }
__device__ inline int mfma_row(float, int k, int r) { return 4 * k + r; }

// Self-test of the lane layout: D = A B for A 16x8, B 8x16 (two MFMA steps)
// against a plain loop.  a, b row-major; out gets D (16x16) from MFMA and ref.
template <typename R, typename V>
__global__ __launch_bounds__(64) void mfma_selftest_kernel(const double *a, const double *b,
                                                           double *out, double *ref) {
  const int l = threadIdx.x, i = l % 16, k = l / 16;
  V acc = {0, 0, 0, 0};
  for (int s = 0; s < 2; s++)
    acc = mfma16x16x4((R)a[i * 8 + 4 * s + k], (R)b[(4 * s + k) * 16 + i], acc);
  for (int r = 0; r < 4; r++) out[mfma_row(R(0), k, r) * 16 + i] = (double)acc[r];
  if (l < 16)
    for (int j = 0; j < 16; j++) {
      double s = 0.0;
      for (int kk = 0; kk < 8; kk++) s += (double)(R)a[l * 8 + kk] * (double)(R)b[kk * 16 + j];
      ref[l * 16 + j] = s;
    }
}

int launch_mfma_selftest(const double *a, const double *b, double *out, double *ref, int fp32,
                         hipStream_t s) {
  if (fp32)
    mfma_selftest_kernel<float, f4_t><<<1, 64, 0, s>>>(a, b, out, ref);
  else
    mfma_selftest_kernel<double, d4_t><<<1, 64, 0, s>>>(a, b, out, ref);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// per-source data of the accurate modes: FP32 (x - c_leaf, y - c_leaf, cir,
// 1/sigma^2) for the mixed form, FP64 (x, y, cir, 1/sigma^2) for the plain one
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void leaf_local_kernel(const NodeRec *__restrict__ nodes,
                                                         const int32_t *__restrict__ leaves,
                                                         int nleaves, const double *__restrict__ sx,
                                                         const double *__restrict__ sy,
                                                         const double *__restrict__ sg,
                                                         const double *__restrict__ ss,
                                                         float4 *__restrict__ loc,
                                                         double4 *__restrict__ loc64) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= nleaves) return;
  const NodeRec r = nodes[leaves[w]];
  const int c = node_count(r.cnt);
  for (int e = lane; e < c; e += 64) {
    const int j = r.first + e;
    const double s2 = ss[j] * ss[j];
    loc64[j] = make_double4(sx[j], sy[j], sg[j], 1.0 / s2);
    loc[j] = make_float4((float)(sx[j] - r.cx), (float)(sy[j] - r.cy), (float)sg[j],
                         (float)(1.0 / s2));
  }
}

int launch_leaf_local(const NodeRec *d_nodes, const int32_t *d_leaves, int nleaves,
                      const double *sx, const double *sy, const double *sg, const double *ss,
                      float4 *loc, double4 *loc64, hipStream_t s) {
  if (nleaves <= 0) return 0;
  leaf_local_kernel<<<(nleaves + 3) / 4, 256, 0, s>>>(d_nodes, d_leaves, nleaves, sx, sy, sg, ss,
                                                      loc, loc64);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// leaves in blob (Morton) order: a leaf starting at blob i marks position i
__global__ __launch_bounds__(256) void mark_leaves_kernel(const NodeRec *__restrict__ nodes,
                                                          int nnodes, int32_t *__restrict__ at,
                                                          uint32_t *__restrict__ flag) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < 4 || i >= nnodes) return;
  const NodeRec r = nodes[i];
  if (node_internal(r.cnt) || node_count(r.cnt) == 0) return;
  at[r.first] = i;
  flag[r.first] = 1u;
}

__global__ __launch_bounds__(256) void compact_leaves_kernel(const int32_t *__restrict__ at,
                                                             const uint32_t *__restrict__ flag,
                                                             const uint32_t *__restrict__ pos,
                                                             int n, int32_t *__restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = at[i];
}

int launch_leaf_list(const NodeRec *d_nodes, int nnodes, int n, int32_t *d_at, uint32_t *d_flag,
                     uint32_t *d_pos, int32_t *d_leaves, uint32_t *d_count, hipStream_t s) {
  FMM_HIP_TRY(hipMemsetAsync(d_flag, 0, sizeof(uint32_t) * n, s));
  mark_leaves_kernel<<<(nnodes + 255) / 256, 256, 0, s>>>(d_nodes, nnodes, d_at, d_flag);
  FMM_HIP_TRY(hipGetLastError());
  if (scan_u32(d_flag, d_pos, n, d_count, s)) return -1;
  compact_leaves_kernel<<<(n + 255) / 256, 256, 0, s>>>(d_at, d_flag, d_pos, n, d_leaves);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

__global__ __launch_bounds__(256) void max_reduce_kernel(const double *__restrict__ v, int n,
                                                         double *__restrict__ out) {
  double m = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = fmax(m, v[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
    // values are >= 0: their bit patterns order like unsigned integers
    atomicMax(reinterpret_cast<unsigned long long *>(out), (unsigned long long)__double_as_longlong(m));
  }
}

int launch_max_reduce(const double *d_v, int n, double *d_out, hipStream_t s) {
  FMM_HIP_TRY(hipMemsetAsync(d_out, 0, sizeof(double), s));
  int nb = (n + 255) / 256;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  max_reduce_kernel<<<nb, 256, 0, s>>>(d_v, n, d_out);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// the fused kernel
// ---------------------------------------------------------------------------
struct Slots2 {
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

template <bool MIXED>
struct FarAcc;

template <>
struct FarAcc<false> {
  d4_t g0 = {0, 0, 0, 0}, g1 = {0, 0, 0, 0};
  __device__ inline void add(double a, double b0, double b1) {
    g0 = mfma16x16x4(a, b0, g0);
    g1 = mfma16x16x4(a, b1, g1);
  }
  __device__ inline void tick() {}
  __device__ inline void fold() {}
  __device__ inline double r0(int r) const { return g0[r]; }
  __device__ inline double r1(int r) const { return g1[r]; }
  __device__ inline int row(int k, int r) const { return mfma_row(0.0, k, r); }
};

template <>
struct FarAcc<true> {
  f4_t c0 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
  d4_t g0 = {0, 0, 0, 0}, g1 = {0, 0, 0, 0};
  int n = 0;
  __device__ inline void add(double a, double b0, double b1) {
    c0 = mfma16x16x4((float)a, (float)b0, c0);
    c1 = mfma16x16x4((float)a, (float)b1, c1);
  }
  __device__ inline void tick() {
    if (++n == 4) fold();
  }
  __device__ inline void fold() {
    for (int r = 0; r < 4; r++) {
      g0[r] += (double)c0[r];
      g1[r] += (double)c1[r];
    }
    c0 = f4_t{0, 0, 0, 0};
    c1 = f4_t{0, 0, 0, 0};
    n = 0;
  }
  __device__ inline double r0(int r) const { return g0[r]; }
  __device__ inline double r1(int r) const { return g1[r]; }
  __device__ inline int row(int k, int r) const { return mfma_row(0.0f, k, r); }
};

// Root of the tree a node belongs to: 0 for a single tree; in a forest the
// system id sits above the quadrant digits of the sorted keys.
__device__ inline int root_of(const FmmArgs &a, int first) {
  return a.root_shift < 0 ? 0 : (int)(a.key[first] >> a.root_shift);
}

// All work of one target leaf (index li of the leaf list) for one wave.
template <bool MIXED>
__device__ void fmm_process_leaf(const FmmArgs &a, int li, int w, int lane,
                                 double2 (*s_loc)[kFmmMaxP], double2 (*s_red)[64],
                                 double4 (*s_stage)[64]) {
  const int A = a.leaves[li];
  const NodeRec NA = a.nodes[A];
  const int cA = node_count(NA.cnt);
  const double RA = NA.region * 0.70710678118654752;
  const double sA = NA.region;
  // this leaf's expansion order and near-field radius
  int pA = a.p;
  double rcut = a.rcut;
  if (a.leaf_par) {
    const float2 lp = a.leaf_par[li];
    pA = min(a.p, max(2, (int)lp.x));
    rcut = (double)lp.y * a.smax;
  }
  // lanes: (target ti, replica rep)
  const int nrep = max(1, 64 / cA);
  const int ti = lane % cA, rep = lane / cA;
  const bool tact = rep < nrep;
  const int tj = NA.first + ti;
  const double tx = a.sx[tj], ty = a.sy[tj];
  float lx = 0.f, ly = 0.f;
  if (MIXED) {
    const float4 q = a.loc[tj];
    lx = q.x;
    ly = q.y;
  }
  double ax = 0.0, ay = 0.0;
  // MFMA lane roles: row / column index i, K slot k
  const int mi = lane & 15, mk = lane >> 4;
  const bool row_ok = mi < pA;
  const double2 sApow = cpow_u(make_double2(sA, 0.0), mi + 1);  // s_A^(i+1), real
  FarAcc<MIXED> acc;
  int pending = -1;

  // one far cell pair (c0, c1; c1 may be -1) -> two MFMA steps
  auto far_pair = [&](int c0, int c1) {
    const int cell = mk < 2 ? c0 : c1;
    double av = 0.0, b0 = 0.0, b1 = 0.0;
    if (cell >= 0 && row_ok) {
      const NodeRec C = a.nodes[cell];
      const double dr = C.cx - NA.cx, di = C.cy - NA.cy;
      const double inv = 1.0 / (dr * dr + di * di);
      const double2 wv = make_double2(dr * inv, -di * inv);
      const double2 wi = cpow_u(wv, mi);                         // w^i
      const double2 pw = cmul(cmul(wi, wv), make_double2(sApow.x, 0.0));  // (s_A w)^(i+1)
      const double2 ak = a.coef[(size_t)cell * a.p + mi];
      double2 q = cmul(ak, wi);                                  // a_i w^i
      if ((mi & 1) == 0) q = make_double2(-q.x, -q.y);           // (-1)^(i+1)
      const bool im = (mk & 1) != 0;
      av = im ? pw.y : pw.x;
      b0 = im ? -q.y : q.x;
      b1 = im ? q.x : q.y;
    }
    acc.add(av, b0, b1);
    acc.tick();
  };

  // Direct sums against one near leaf.  With one replica (every lane needs
  // the same source at each step) the source index is uniform and the
  // sources come through the scalar cache; otherwise the leaf's sources are
  // staged once in this wave's LDS slice and read from there, so each source
  // crosses the vector memory path once per wave instead of once per lane.
  auto near_leaf = [&](int c) {
    const NodeRec C = a.nodes[c];
    const int cnt = node_count(C.cnt);
    const int f = __builtin_amdgcn_readfirstlane(C.first);
    if (MIXED) {
      const float dx = (float)(C.cx - NA.cx), dy = (float)(C.cy - NA.cy);
      const float px = lx - dx, py = ly - dy;
      float sxs = 0.f, sxc = 0.f, sys = 0.f, syc = 0.f;
      // pair (target, source s): u += -f ry, v += f rx with
      // f = cir (1 - exp(-r^2/sigma^2)) / r^2, pairs closer than 1e-10 skipped
      auto pair = [&](const float4 s, float &gx, float &gy) {
        const float rx = px - s.x, ry = py - s.y;
        const float r2 = fmaf(rx, rx, ry * ry);
        float fac = s.z * one_minus_exp_neg(r2 * s.w) * __builtin_amdgcn_rcpf(r2);
        fac = r2 > 1e-20f ? fac : 0.0f;
        gx = fmaf(-fac, ry, gx);
        gy = fmaf(fac, rx, gy);
      };
      // groups of four sources: plain FP32 sum inside a group, compensated
      // (Kahan) accumulation of the group sums
      if (nrep == 1) {
        const float4 *src = a.loc + f;
        int j = 0;
        for (; j + 3 < cnt; j += 4) {
          const float4 s0 = src[j], s1 = src[j + 1], s2 = src[j + 2], s3 = src[j + 3];
          float gx = 0.f, gy = 0.f;
          pair(s0, gx, gy);
          pair(s1, gx, gy);
          pair(s2, gx, gy);
          pair(s3, gx, gy);
          kahan_add(sxs, sxc, gx);
          kahan_add(sys, syc, gy);
        }
        float gx = 0.f, gy = 0.f;
        for (; j < cnt; j++) pair(src[j], gx, gy);
        kahan_add(sxs, sxc, gx);
        kahan_add(sys, syc, gy);
      } else {
        float4 *st = reinterpret_cast<float4 *>(s_stage[w]);
        if (lane < cnt) st[lane] = a.loc[f + lane];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (tact) {
          int j = rep;
          for (; j + 3 * nrep < cnt; j += 4 * nrep) {
            const float4 s0 = st[j], s1 = st[j + nrep], s2 = st[j + 2 * nrep],
                         s3 = st[j + 3 * nrep];
            float gx = 0.f, gy = 0.f;
            pair(s0, gx, gy);
            pair(s1, gx, gy);
            pair(s2, gx, gy);
            pair(s3, gx, gy);
            kahan_add(sxs, sxc, gx);
            kahan_add(sys, syc, gy);
          }
          float gx = 0.f, gy = 0.f;
          for (; j < cnt; j += nrep) pair(st[j], gx, gy);
          kahan_add(sxs, sxc, gx);
          kahan_add(sys, syc, gy);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      ax += (double)sxs - (double)sxc;
      ay += (double)sys - (double)syc;
    } else {
      auto pair = [&](const double4 s) {
        const double rx = tx - s.x, ry = ty - s.y;
        const double r2 = fma(rx, rx, ry * ry);
        // 1 / r^2: hardware reciprocal refined by two Newton steps
        double ri = __builtin_amdgcn_rcp(r2);
        ri = ri * fma(-r2, ri, 2.0);
        ri = ri * fma(-r2, ri, 2.0);
        double fac = s.z * (1.0 - exp(-r2 * s.w)) * ri;
        fac = r2 > 1e-20 ? fac : 0.0;
        ax = fma(-fac, ry, ax);
        ay = fma(fac, rx, ay);
      };
      if (nrep == 1) {
        const double4 *src = a.loc64 + f;
        int j = 0;
        for (; j + 3 < cnt; j += 4) {
          const double4 s0 = src[j], s1 = src[j + 1], s2 = src[j + 2], s3 = src[j + 3];
          pair(s0);
          pair(s1);
          pair(s2);
          pair(s3);
        }
        for (; j < cnt; j++) pair(src[j]);
      } else {
        double4 *st = s_stage[w];
        if (lane < cnt) st[lane] = a.loc64[f + lane];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (tact) {
          int j = rep;
          for (; j + 3 * nrep < cnt; j += 4 * nrep) {
            const double4 s0 = st[j], s1 = st[j + nrep], s2 = st[j + 2 * nrep],
                          s3 = st[j + 3 * nrep];
            pair(s0);
            pair(s1);
            pair(s2);
            pair(s3);
          }
          for (; j < cnt; j += nrep) pair(st[j]);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
  };

  // wave-uniform traversal of the whole tree (masks of opened children per level)
  // (node indices are uniform across the wave: readfirstlane lets the node
  // records come through scalar loads)
  Slots2 sl;
  int d = 0, cur = 0, node = __builtin_amdgcn_readfirstlane(root_of(a, NA.first));
  bool expand = true;
  for (;;) {
    if (expand) {
      const int base = __builtin_amdgcn_readfirstlane(a.nodes[node].first);
      uint32_t m = 0;
      for (int q = 0; q < 4; q++) {
        const int c = base + q;
        const NodeRec C = a.nodes[c];
        if (node_count(C.cnt) == 0) continue;
        const double dx = C.cx - NA.cx, dy = C.cy - NA.cy;
        const double dd = sqrt(dx * dx + dy * dy);
        const double RC = C.region * 0.70710678118654752;
        if (RA + RC <= a.theta * dd && dd - RA - RC >= rcut) {
          if (pending < 0) {
            pending = c;
          } else {
            far_pair(pending, c);
            pending = -1;
          }
        } else {
          m |= 1u << q;
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
      cur = __builtin_amdgcn_readfirstlane(a.gparent[cur >> 2]);
      d--;
    }
    if (done) break;
    const int q = 31 - __clz(m);
    sl.set(d, m & ~(1u << q));
    node = __builtin_amdgcn_readfirstlane((cur & ~3) + q);
    const int32_t cn = a.nodes[node].cnt;
    if (node_internal(cn)) {
      expand = true;
    } else {
      near_leaf(node);
      expand = false;
    }
  }
  if (pending >= 0) far_pair(pending, -1);
  acc.fold();

  // local expansion: Lt_l = (1/s_A) sum_k C(l+k,k) G[l][k]; lane holds rows
  // acc.row(mk, r), column mi of G (real part in r0, imaginary in r1)
  for (int r = 0; r < 4; r++) {
    const int row = acc.row(mk, r);
    // C(row + mi, mi)
    double bc = 1.0;
    for (int t = 1; t <= mi; t++) bc = bc * (double)(row + t) / (double)t;
    double gr = (mi < pA) ? acc.r0(r) * bc : 0.0;
    double gi = (mi < pA) ? acc.r1(r) * bc : 0.0;
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
      gr += __shfl_xor(gr, o, 64);
      gi += __shfl_xor(gi, o, 64);
    }
    if (mi == 0 && row < kFmmMaxP) s_loc[w][row] = make_double2(gr / sA, gi / sA);
  }
  // s_loc[w] and s_red[w] belong to this wave alone: a wave-level fence orders
  // the LDS writes of some lanes before the reads of others
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double phr = 0.0, phi = 0.0;
  if (tact && rep == 0) {
    // L2P: sum_l Lt_l u^l, u = (z - c_A)/s_A
    const double ur = (tx - NA.cx) / sA, ui = (ty - NA.cy) / sA;
    double2 s = s_loc[w][pA - 1];
    for (int l = pA - 2; l >= 0; l--) {
      const double2 L = s_loc[w][l];
      s = make_double2(s.x * ur - s.y * ui + L.x, s.x * ui + s.y * ur + L.y);
    }
    phr = s.x;
    phi = s.y;
  }
  s_red[w][lane] = make_double2(ax, ay);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (tact && rep == 0) {
    double sxa = 0.0, sya = 0.0;
    for (int rr = 0; rr < nrep; rr++) {
      const double2 v = s_red[w][rr * cA + ti];
      sxa += v.x;
      sya += v.y;
    }
    const double u = 0.5 * phi / M_PI + 0.5 * sxa / M_PI;
    const double v = 0.5 * phr / M_PI + 0.5 * sya / M_PI;
    a.uv[a.perm[tj]] = make_double2(u, v);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One wave per target leaf, leaf = 4 * block + wave.
template <bool MIXED>
__global__ __launch_bounds__(256) void fmm_leaf_kernel(FmmArgs a) {
  __shared__ double2 s_loc[4][kFmmMaxP];
  __shared__ double2 s_red[4][64];
  __shared__ double4 s_stage[4][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = blockIdx.x * 4 + w;
  if (li >= a.nleaves) return;
  fmm_process_leaf<MIXED>(a, li, w, lane, s_loc, s_red, s_stage);
}

// Persistent form: a fixed grid whose waves take target leaves from a work
// counter until the device-side leaf count is exhausted.  Every wave exits
// once the counter passes the count, so the grid always drains.  Nothing in
// the launch depends on host-side counts, so it can be captured in a graph
// and replayed on a rebuilt tree.
template <bool MIXED>
__global__ __launch_bounds__(256) void fmm_leaf_persistent_kernel(FmmArgs a,
                                                                  const uint32_t *__restrict__ nleaves,
                                                                  uint32_t *__restrict__ counter) {
  __shared__ double2 s_loc[4][kFmmMaxP];
  __shared__ double2 s_red[4][64];
  __shared__ double4 s_stage[4][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t total = *nleaves;
  for (;;) {
    uint32_t li = 0;
    if (lane == 0) li = atomicAdd(counter, 1u);
    li = __shfl(li, 0, 64);
    if (li >= total) break;
    fmm_process_leaf<MIXED>(a, (int)li, w, lane, s_loc, s_red, s_stage);
  }
}

__global__ void zero_counter_kernel(uint32_t *c) { *c = 0u; }

int launch_fmm_eval_persistent(const FmmArgs &args, bool mixed, const uint32_t *d_nleaves,
                               uint32_t *d_counter, int nblocks, hipStream_t s) {
  zero_counter_kernel<<<1, 1, 0, s>>>(d_counter);
  FMM_HIP_TRY(hipGetLastError());
  if (mixed)
    fmm_leaf_persistent_kernel<true><<<nblocks, 256, 0, s>>>(args, d_nleaves, d_counter);
  else
    fmm_leaf_persistent_kernel<false><<<nblocks, 256, 0, s>>>(args, d_nleaves, d_counter);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// Work of the fused kernel, one thread per target leaf with the same walk:
// out[2*leaf] = far cells (each one M2L into the leaf), out[2*leaf+1] = near
// source-target pairs.
__global__ __launch_bounds__(256) void fmm_count_kernel(FmmArgs a, int64_t *__restrict__ out) {
  const int li = blockIdx.x * 256 + threadIdx.x;
  if (li >= a.nleaves) return;
  const NodeRec NA = a.nodes[a.leaves[li]];
  const int cA = node_count(NA.cnt);
  const double RA = NA.region * 0.70710678118654752;
  const double rcut = a.leaf_par ? (double)a.leaf_par[li].y * a.smax : a.rcut;
  int64_t far = 0, pairs = 0;
  Slots2 sl;
  int d = 0, cur = 0, node = root_of(a, NA.first);
  bool expand = true;
  for (;;) {
    if (expand) {
      const int base = a.nodes[node].first;
      uint32_t m = 0;
      for (int q = 0; q < 4; q++) {
        const NodeRec C = a.nodes[base + q];
        if (node_count(C.cnt) == 0) continue;
        const double dx = C.cx - NA.cx, dy = C.cy - NA.cy;
        const double dd = sqrt(dx * dx + dy * dy);
        const double RC = C.region * 0.70710678118654752;
        if (RA + RC <= a.theta * dd && dd - RA - RC >= rcut) far++;
        else m |= 1u << q;
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
      cur = a.gparent[cur >> 2];
      d--;
    }
    if (done) break;
    const int q = 31 - __clz(m);
    sl.set(d, m & ~(1u << q));
    node = (cur & ~3) + q;
    const int32_t cn = a.nodes[node].cnt;
    if (node_internal(cn)) {
      expand = true;
    } else {
      pairs += (int64_t)cA * node_count(cn);
      expand = false;
    }
  }
  out[2 * li] = far;
  out[2 * li + 1] = pairs;
}

int launch_fmm_count(const FmmArgs &args, int64_t *d_out, hipStream_t s) {
  if (args.nleaves <= 0) return 0;
  fmm_count_kernel<<<(args.nleaves + 255) / 256, 256, 0, s>>>(args, d_out);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_fmm_eval(const FmmArgs &args, bool mixed, hipStream_t s) {
  if (args.nleaves <= 0) return 0;
  const int nb = (args.nleaves + 3) / 4;
  if (mixed)
    fmm_leaf_kernel<true><<<nb, 256, 0, s>>>(args);
  else
    fmm_leaf_kernel<false><<<nb, 256, 0, s>>>(args);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace fmm2d
