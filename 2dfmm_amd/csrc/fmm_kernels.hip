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
//     sum_l Lt_l ((z - c_A)/s_A)^l (L2P);
//   * L2L: cells far from every target-leaf child of a parent P are
//     contracted once into P's local expansion (parent_local_kernel), which
//     each child starts from, shifted to its own centre; the child's walk
//     skips them;
//   * smooth cells (mixed form, one core radius): near nodes of the smooth
//     level are summed for every target through a Taylor expansion of the
//     regularised kernel about the cell centre (smooth_eval).
// The traversal decisions are uniform across the wave (they depend on the
// target leaf only), so the wave walks the tree together; lanes are
// (target, replica) pairs for the direct part.
//
// Precision: FP64 everywhere for the plain form.  The mixed form runs the
// direct part in FP32 on coordinates relative to the target leaf's centre
// (sums of eight pairs in FP32, accumulated in FP64), the smooth cells in
// FP32 (each value added to the FP64 sums), and the contraction on FP32 MFMA
// (promoted to FP64 every 8 cells); expansions (S2M/M2M), L2L and the L2P
// accumulation stay in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fmm2d_internal.h"

namespace fmm2d {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

constexpr int kFmmMaxP = 16;

// C(r + i, i) for r, i < kFmmMaxP (the M2L weights of a far-cell burst):
// integers below 2^28, so every entry is exact in FP64.
struct M2lBinom {
  double v[kFmmMaxP][kFmmMaxP];
};
constexpr M2lBinom make_m2l_binom() {
  M2lBinom t{};
  for (int r = 0; r < kFmmMaxP; r++) {
    double b = 1.0;
    for (int i = 0; i < kFmmMaxP; i++) {
      if (i > 0) b = b * (double)(r + i) / (double)i;
      t.v[r][i] = b;
    }
  }
  return t;
}
__constant__ M2lBinom kM2lBinom = make_m2l_binom();

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// a b with fused multiply-adds
__device__ inline double2 cmul_fma(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}

// w^i for 0 <= i < 16 in the same steps in every lane (a loop over the bits
// of a lane's own exponent runs the longest trip in every lane): the product
// of w^(2^b) over the set bits b of i.
__device__ inline double2 cpow16(double2 w, int i) {
  const double2 one = make_double2(1.0, 0.0);
  const double2 w2 = cmul_fma(w, w), w4 = cmul_fma(w2, w2), w8 = cmul_fma(w4, w4);
  const double2 lo = cmul_fma((i & 1) ? w : one, (i & 2) ? w2 : one);
  const double2 hi = cmul_fma((i & 4) ? w4 : one, (i & 8) ? w8 : one);
  return cmul_fma(lo, hi);
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

// Pair kernel of the mixed form.  Each source stores w = log2(e) / sigma^2
// (capped at 7.2e19) and z = cir w, so with y = r^2 w
//   f = cir (1 - exp(-r^2 / sigma^2)) / r^2 = z h(y),  h(y) = (1 - 2^-y) / y.
// h is (1 - 2^-y) / y for y >= 0.7213 (2^-y <= 0.61: no cancellation; one
// v_exp_f32 and one v_rcp_f32) and its Taylor series below, which a wave
// runs only when one of its lanes needs it.  Pairs closer than 1e-10 (the
// direct sum's cut-off, r^2 <= 1e-20) have y < 0.7213 because of the cap
// and are zeroed inside the series branch; the cap makes a zero core radius
// (mirror images) the point vortex for every pair beyond that cut-off up to
// a 2^-y term, y >= r^2 7.2e19.
constexpr float kSeriesY = 0.7213475f;  // y ln 2 = 0.5
constexpr double kMaxW = 7.2e19;

// h(y) = sum_k ln2 (-ln2)^k y^k / (k+1)!, k <= 7; truncation below 1.1e-8
// relative on [0, 0.7213].  Explicit fused multiply-adds (the build
// disables contraction).
__device__ inline float h_series(float y) {
  float q = fmaf(y, -1.3215487e-06f, 1.5252734e-05f);
  q = fmaf(y, q, -1.5403530e-04f);
  q = fmaf(y, q, 1.3333558e-03f);
  q = fmaf(y, q, -9.6181291e-03f);
  q = fmaf(y, q, 5.5504109e-02f);
  q = fmaf(y, q, -2.4022651e-01f);
  return fmaf(y, q, 6.9314718e-01f);
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
// per-source data of the accurate modes: FP32 (x - c_leaf, y - c_leaf, z, w)
// with w = log2(e)/sigma^2 (capped) and z = cir w for the mixed form, FP64
// (x, y, cir, 1/sigma^2) for the plain one
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void leaf_local_kernel(const NodeRec *__restrict__ nodes,
                                                         const int32_t *__restrict__ leaves,
                                                         const uint32_t *__restrict__ count,
                                                         const double *__restrict__ sx,
                                                         const double *__restrict__ sy,
                                                         const double *__restrict__ sg,
                                                         const double *__restrict__ ss,
                                                         float4 *__restrict__ loc,
                                                         double4 *__restrict__ loc64,
                                                         const double *__restrict__ smax) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= (int)*count) return;
  const double sw = smax[3];
  const NodeRec r = nodes[leaves[w]];
  const int c = node_count(r.cnt);
  for (int e = lane; e < c; e += 64) {
    const int j = r.first + e;
    const double s2 = ss[j] * ss[j];
    loc64[j] = make_double4(sx[j], sy[j], sg[j], 1.0 / s2);
    if (sw > 0.0) {
      // one core radius: coordinates times sqrt(w), z = cir sqrt(w), y = r'^2
      loc[j] = make_float4((float)((sx[j] - r.cx) * sw), (float)((sy[j] - r.cy) * sw),
                           (float)(sg[j] * sw), 1.0f);
    } else {
      const double wj = fmin(1.4426950408889634 / s2, kMaxW);
      loc[j] = make_float4((float)(sx[j] - r.cx), (float)(sy[j] - r.cy), (float)(sg[j] * wj),
                           (float)wj);
    }
  }
}

int launch_leaf_local(const NodeRec *d_nodes, const int32_t *d_leaves, const uint32_t *d_count,
                      int bound, const double *sx, const double *sy, const double *sg,
                      const double *ss, float4 *loc, double4 *loc64, const double *d_smax,
                      hipStream_t s) {
  if (bound <= 0) return 0;
  leaf_local_kernel<<<(bound + 3) / 4, 256, 0, s>>>(d_nodes, d_leaves, d_count, sx, sy, sg, ss,
                                                    loc, loc64, d_smax);
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

// A leaf whose box is large against the near radius has a near region of
// about 1/theta times its box (a sparse box of a few blobs next to a dense
// cluster sums pairs with much of the cluster), and its walk takes up to ten
// times a typical leaf's.  Such leaves are started first, the largest first:
// key = 9 + floor(log2(rcut / R)) for R > rcut (half diagonal R), 9 for the
// others, which keep their Morton order (neighbours in space run together).
__global__ __launch_bounds__(256) void leaf_launch_keys_kernel(const NodeRec *__restrict__ nodes,
                                                               const int32_t *__restrict__ leaves,
                                                               int n_bound,
                                                               const uint32_t *__restrict__ count,
                                                               const double *__restrict__ smax,
                                                               double cut,
                                                               uint64_t *__restrict__ key,
                                                               uint32_t *__restrict__ val,
                                                               uint32_t *__restrict__ front) {
  const int li = blockIdx.x * 256 + threadIdx.x;
  if (li >= (int)*count || li >= n_bound) return;
  const double rcut = cut * smax[0];
  const double R = nodes[leaves[li]].region * 0.70710678118654752;
  int k = 9;
  if (R > rcut) k = min(8, max(0, 9 + ilogb(rcut / R)));
  key[li] = (uint64_t)k;
  val[li] = (uint32_t)li;
  const uint64_t fm = __builtin_amdgcn_ballot_w64(k < 9);
  if (fm && (threadIdx.x & 63) == (uint32_t)__builtin_ctzll(fm))
    atomicAdd(front, (uint32_t)__builtin_popcountll(fm));
}

int launch_leaf_launch_keys(const NodeRec *d_nodes, const int32_t *d_leaves, int n_bound,
                            const uint32_t *d_count, const double *d_smax, double cut,
                            uint64_t *d_key, uint32_t *d_val, uint32_t *d_front, hipStream_t s) {
  FMM_HIP_TRY(hipMemsetAsync(d_front, 0, sizeof(uint32_t), s));
  if (n_bound <= 0) return 0;
  leaf_launch_keys_kernel<<<(n_bound + 255) / 256, 256, 0, s>>>(d_nodes, d_leaves, n_bound, d_count,
                                                                d_smax, cut, d_key, d_val, d_front);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// out[0] = max, out[1] = min of n values >= 0 (their bit patterns order like
// unsigned integers, so integer atomics combine the blocks).
__global__ __launch_bounds__(256) void minmax_reduce_kernel(const double *__restrict__ v, int n,
                                                            double *__restrict__ out) {
  double hi = 0.0, lo = 1.0e300;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    hi = fmax(hi, v[i]);
    lo = fmin(lo, v[i]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    lo = fmin(lo, __shfl_xor(lo, o, 64));
  }
  __shared__ double s[2][4];
  if ((threadIdx.x & 63) == 0) {
    s[0][threadIdx.x >> 6] = hi;
    s[1][threadIdx.x >> 6] = lo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    hi = fmax(fmax(s[0][0], s[0][1]), fmax(s[0][2], s[0][3]));
    lo = fmin(fmin(s[1][0], s[1][1]), fmin(s[1][2], s[1][3]));
    atomicMax(reinterpret_cast<unsigned long long *>(out), (unsigned long long)__double_as_longlong(hi));
    atomicMin(reinterpret_cast<unsigned long long *>(out + 1),
              (unsigned long long)__double_as_longlong(lo));
  }
}

__global__ void minmax_init_kernel(double *out) {
  out[0] = 0.0;
  out[1] = 1.0e300;
}

// out[2] = w_max = log2(e) / sigma_min^2 (capped at kMaxW), out[3] = sqrt(w_max)
// with one core radius, else 0: the pair kernel's scales, formed once on the
// device so that every kernel and the host use the same values.
__global__ void minmax_finish_kernel(double *out) {
  const double w = out[1] > 0.0 ? fmin(1.4426950408889634 / (out[1] * out[1]), kMaxW) : kMaxW;
  out[2] = w;
  out[3] = out[0] == out[1] ? sqrt(w) : 0.0;
}

// Child 0 of a node starts at the node's first blob, so following child 0
// down to a leaf gives the subtree's first blob.
__global__ __launch_bounds__(256) void node_first_blob_kernel(const NodeRec *__restrict__ nodes,
                                                              int nnodes,
                                                              int32_t *__restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nnodes) return;
  NodeRec c = nodes[i];
  while (node_internal(c.cnt)) c = nodes[c.first];
  out[i] = c.first;
}

int launch_node_first_blob(const NodeRec *d_nodes, int nnodes, int32_t *d_out, hipStream_t s) {
  if (nnodes <= 0) return 0;
  node_first_blob_kernel<<<(nnodes + 255) / 256, 256, 0, s>>>(d_nodes, nnodes, d_out);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_minmax_reduce(const double *d_v, int n, double *d_out, hipStream_t s) {
  minmax_init_kernel<<<1, 1, 0, s>>>(d_out);
  FMM_HIP_TRY(hipGetLastError());
  int nb = (n + 255) / 256;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  minmax_reduce_kernel<<<nb, 256, 0, s>>>(d_v, n, d_out);
  FMM_HIP_TRY(hipGetLastError());
  minmax_finish_kernel<<<1, 1, 0, s>>>(d_out);
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
  typedef double T;  // type of the MFMA operands
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
  typedef float T;
  __device__ inline void add(float a, float b0, float b1) {
    c0 = mfma16x16x4(a, b0, c0);
    c1 = mfma16x16x4(a, b1, c1);
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

// ---------------------------------------------------------------------------
// Work of one target leaf.  A wave walks the tree for its leaf; far cells go
// into the MFMA contraction, near leaves are summed directly.  Leaves with
// much near work (dense clusters: a few leaves can carry hundreds of
// thousands of near sources) are split: the fused pass sums the first
// `split_budget` work units itself and saves a checkpoint of the walk every
// `split_budget` units after that; continuation chunks resume the walk from
// the checkpoints on other waves, and a last pass adds their partial sums in
// chunk order (deterministic).
// ---------------------------------------------------------------------------
constexpr int kMaxCk = 128;    // chunk boundaries per leaf kept in LDS
constexpr int kFarList = 128;  // far cells listed per burst

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ inline int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Arrays that no kernel of the evaluation writes (tree, sources, leaf list,
// checkpoint records in the continuation pass) are read through the
// constant address space: loads at wave-uniform indices then go through the
// scalar unit even where the kernel stores elsewhere.
template <typename T>
__device__ inline T ldro(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized records only");
  typedef const uint32_t __attribute__((address_space(4))) *cwords;
  const cwords q = (cwords)p;
  T v;
  uint32_t *d = reinterpret_cast<uint32_t *>(&v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); i++) d[i] = q[i];
  return v;
}
__device__ inline uint64_t rfl64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// The value of lane b (wave-uniform) in every lane: v_readlane, no LDS
// round trip.
__device__ inline int lane_of(int v, int b) { return __builtin_amdgcn_readlane(v, b); }
__device__ inline double lane_of(double v, int b) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, b);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), b);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// the record of node held by lane b
__device__ inline NodeRec node_of_lane(const NodeRec &N, int b) {
  NodeRec S;
  S.cx = lane_of(N.cx, b);
  S.cy = lane_of(N.cy, b);
  S.region = lane_of(N.region, b);
  S.first = lane_of(N.first, b);
  S.cnt = lane_of(N.cnt, b);
  return S;
}

// LDS stage of one wave: up to 64 sources of a near step, replica-major
// (replica r reads its sources contiguously from r * stride), so rows of
// up to 128 entries; FP32 entries for the mixed form, FP64 for the plain one.
template <bool MIXED>
struct StageOf {
  typedef double4 T;
};
template <>
struct StageOf<true> {
  typedef float4 T;
};
constexpr int kStageSlots = 128;
// The mixed form's stage has room past kStageSlots for the coefficients of
// one smooth batch (3 cells), their offsets and the 64 targets' coordinates.
constexpr int kSmoothResSlots = 192;
template <bool MIXED>
constexpr int stage_slots() { return MIXED ? kStageSlots + kSmoothResSlots : kStageSlots; }

// Per-target-leaf values; lanes are (target ti, replica rep) pairs.
// Source j of a staged step goes to LDS slot (j mod nrep) stride + j / nrep
// (spos for j = lane), stride = ceil(64 / nrep).
// Smooth cells are summed up to kSmoothBatch at a time (the stage holds
// their coefficients).  A batch of nb cells is nb x cA evaluations, (target,
// cell) pairs dealt to the 64 lanes in rounds; nb is chosen to fill the
// rounds best, the larger on a tie.
constexpr int kSmoothBatch = 3;
__device__ inline int smooth_batch_size(int cA) {
  int best = 1, bu = 0;  // bu: evaluations per 64 lane slots, x 1024
  for (int nb = 1; nb <= kSmoothBatch; nb++) {
    const int ne = nb * cA;
    const int u = (ne * 1024) / (64 * ((ne + 63) >> 6));
    if (u >= bu) {
      bu = u;
      best = nb;
    }
  }
  return best;
}

struct LeafCtx {
  NodeRec NA;
  int cA, pA, nrep, ti, rep, tj, stride, spos;
  int sbsz;       // smooth batch size
  uint32_t rinv;  // ceil(2^20 / nrep): divisions by nrep as a multiply
  int64_t su;     // work units of one smooth cell
  bool tact;
  double RA, sA, rcut, inv_theta, tx, ty;
  float lx, ly;
};

template <bool MIXED, bool UNI>
__device__ inline LeafCtx leaf_ctx(const FmmArgs &a, int li, int lane) {
  LeafCtx L;
  L.NA = ldro(a.nodes + rfl(ldro(a.leaves + li)));
  L.cA = node_count(L.NA.cnt);
  L.RA = L.NA.region * 0.70710678118654752;
  L.sA = L.NA.region;
  L.inv_theta = 1.0 / a.theta;
  // this leaf's expansion order and near-field radius
  L.pA = a.p;
  L.rcut = a.rcut;
  if (a.leaf_par) {
    const float2 lp = a.leaf_par[li];
    L.pA = min(a.p, max(2, (int)lp.x));
    L.rcut = (double)lp.y * a.smax;
  }
  L.nrep = max(1, 64 / L.cA);
  L.rinv = ((1u << 20) + (uint32_t)L.nrep - 1u) / (uint32_t)L.nrep;
  L.sbsz = smooth_batch_size(L.cA);
  // a cell's share of its batch's time: the batch's rounds over its cells
  L.su = (kSmoothUnits * ((L.sbsz * L.cA + 63) >> 6) + L.sbsz - 1) / L.sbsz;
  L.ti = lane % L.cA;
  L.rep = lane / L.cA;
  L.stride = (64 + L.nrep - 1) / L.nrep;
  L.spos = (lane % L.nrep) * L.stride + lane / L.nrep;
  L.tact = L.rep < L.nrep;
  L.tj = L.NA.first + L.ti;
  // FP64 target coordinates: the plain form's pairs; the mixed form needs
  // them only for the L2P at the end and loads them there (nothing to keep
  // live through the walk)
  L.tx = MIXED ? 0.0 : a.sx[L.tj];
  L.ty = MIXED ? 0.0 : a.sy[L.tj];
  L.lx = 0.f;
  L.ly = 0.f;
  if (MIXED) {
    const float4 q = a.loc[L.tj];
    L.lx = q.x;
    L.ly = q.y;
  }
  return L;
}

// Work units of one near step over cnt sources: wave steps plus a fixed
// cost for the step's loads and synchronisation.
__device__ inline int64_t near_units(int cnt, const LeafCtx &L) {
  // ceil(cnt / nrep), exact for cnt + nrep < 2^12
  return (int64_t)(((uint64_t)(cnt + L.nrep - 1) * L.rinv) >> 20) + kStepUnits;
}

// One pair (target, source s) of the mixed form, target at (px, py) and
// source at (s.x, s.y) in the same FP32 frame: u += -f ry, v += f rx with
// f = z h(y), y = r^2 s.w (see h_series).  SERIES = false for near items
// whose pairs all have y >= kSeriesY (needs_series): no check at all.
template <bool SERIES, bool UNI>
__device__ inline void pair32(float px, float py, const float4 s, float &gx, float &gy) {
  const float rx = px - s.x, ry = py - s.y;
  const float r2 = fmaf(rx, rx, ry * ry);
  const float y = UNI ? r2 : r2 * s.w;
  // z (1 - 2^-y) / y with z - z 2^-y in one fused multiply-add
  float fac = fmaf(-s.z, __builtin_amdgcn_exp2f(-y), s.z) * __builtin_amdgcn_rcpf(y);
  if (SERIES) {
    if (__builtin_amdgcn_fcmpf(y, kSeriesY, 4 /* ordered less-than, wave mask */)) {
      if (y < kSeriesY) fac = r2 > 1e-20f ? s.z * h_series(y) : 0.0f;
    }
  }
  gx = fmaf(-fac, ry, gx);
  gy = fmaf(fac, rx, gy);
}

// One pair of the plain form, FP64 (x, y, cir, 1/sigma^2); the direct sum's
// rule: pairs closer than 1e-10 are skipped.
__device__ inline void pair64(double tx, double ty, const double4 s, double &ax, double &ay) {
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
}

// The sources of one replica: source j = rep + k nrep (< cnt) at row[k] of
// the replica-major stage.  Groups of eight pairs are summed in FP32 and
// accumulated in FP64; the row is read with immediate offsets.
template <bool SERIES, bool UNI>
__device__ inline void lds_sum32(float px, float py, const float4 *row, int cnt, int rep, int nrep,
                                 double &ax, double &ay) {
  int j = rep;
  for (; j + 7 * nrep < cnt; j += 8 * nrep, row += 8) {
    // all eight reads issued before the first pair (the series branch would
    // otherwise keep them apart)
    float4 v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = row[k];
    float gx = 0.f, gy = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) pair32<SERIES, UNI>(px, py, v[k], gx, gy);
    ax += (double)gx;
    ay += (double)gy;
  }
  if (j < cnt) {
    float gx = 0.f, gy = 0.f;
    for (; j < cnt; j += nrep) pair32<SERIES, UNI>(px, py, *row++, gx, gy);
    ax += (double)gx;
    ay += (double)gy;
  }
}

__device__ inline void lds_sum64(double tx, double ty, const double4 *row, int cnt, int rep,
                                 int nrep, double &ax, double &ay) {
  int j = rep;
  for (; j + 3 * nrep < cnt; j += 4 * nrep, row += 4) {
#pragma unroll
    for (int k = 0; k < 4; k++) pair64(tx, ty, row[k], ax, ay);
  }
  for (; j < cnt; j += nrep) pair64(tx, ty, *row++, ax, ay);
}

// Whether a pair of target leaf A and near item C can have y < kSeriesY:
// y = r^2 w >= gap^2 wmax is false only if the gap between the two boxes
// (half side region/2) is below sqrt(kSeriesY / wmax); 1% margin for the FP32
// coordinates.  With one core radius this is exact up to the margin.
__device__ inline bool needs_series(const FmmArgs &a, const LeafCtx &L, const NodeRec &C) {
  const double h = 0.5 * (L.NA.region + C.region);
  const double gx = fmax(fabs(C.cx - L.NA.cx) - h, 0.0), gy = fmax(fabs(C.cy - L.NA.cy) - h, 0.0);
  return (gx * gx + gy * gy) * (double)a.wmax < 1.01 * (double)kSeriesY;
}

// Direct sums of the targets of L against near leaf C.  The leaf's sources
// are staged once in this wave's LDS slice `st` (one vector load per lane)
// and read from there, so each source crosses the vector memory path once
// per wave; with one replica every lane reads the same LDS word (a
// broadcast).  For the mixed form this measured 5% faster than reading the
// sources at wave-uniform addresses through the scalar cache, four loads at
// a time, whose latency the pair arithmetic did not cover; the plain form
// measured the same either way.
template <bool MIXED, bool UNI>
__device__ inline void near_leaf(const FmmArgs &a, const LeafCtx &L, const NodeRec &C, bool ser,
                                 int lane, typename StageOf<MIXED>::T *st, double &ax,
                                 double &ay) {
  const int cnt = node_count(C.cnt);
  const int f = rfl(C.first);
  const int nrep = L.nrep;
  if (MIXED) {
    // sources are stored relative to their own leaf's centre (times sqrt(w)
    // with one core radius)
    const float dx = UNI ? (float)((C.cx - L.NA.cx) * a.sw) : (float)(C.cx - L.NA.cx);
    const float dy = UNI ? (float)((C.cy - L.NA.cy) * a.sw) : (float)(C.cy - L.NA.cy);
    const float px = L.lx - dx, py = L.ly - dy;
    float4 *sf = reinterpret_cast<float4 *>(st);
    if (lane < cnt) sf[L.spos] = a.loc[f + lane];
    wave_sync();
    if (L.tact) {
      if (ser)
        lds_sum32<true, UNI>(px, py, sf + L.rep * L.stride, cnt, L.rep, nrep, ax, ay);
      else
        lds_sum32<false, UNI>(px, py, sf + L.rep * L.stride, cnt, L.rep, nrep, ax, ay);
    }
    wave_sync();
  } else {
    double4 *sd = reinterpret_cast<double4 *>(st);
    if (lane < cnt) sd[L.spos] = a.loc64[f + lane];
    wave_sync();
    if (L.tact) lds_sum64(L.tx, L.ty, sd + L.rep * L.stride, cnt, L.rep, nrep, ax, ay);
    wave_sync();
  }
}

// Direct sums of the targets of L against tiles ta..tb-1 (64 sources each)
// of a contiguous run of cnt sources from s0 spanning several leaves (a whole
// near subtree, or consecutive near leaves of one parent).  Each tile is staged in LDS -- the mixed form converts each source
// once, to FP32 coordinates relative to the target leaf's centre -- and the
// next tile's loads are issued before the current tile's sums.
template <bool MIXED, bool UNI>
__device__ inline void near_range(const FmmArgs &a, const LeafCtx &L, int s0, int cnt, bool ser,
                                  int ta, int tb, int lane, typename StageOf<MIXED>::T *st,
                                  double &ax, double &ay) {
  double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
  float2 zw = make_float2(0.f, 0.f);
  auto load = [&](int t) {
    const int i = 64 * t + lane;
    if (i < cnt) {
      q = a.loc64[s0 + i];
      if (MIXED) {
        const float4 p = a.loc[s0 + i];
        zw = make_float2(p.z, p.w);
      }
    }
  };
  load(ta);
  for (int t = ta; t < tb; t++) {
    const int tc = min(64, cnt - 64 * t);
    if (MIXED) {
      float4 *sf = reinterpret_cast<float4 *>(st);
      if (lane < tc)
        sf[L.spos] = UNI ? make_float4((float)((q.x - L.NA.cx) * a.sw),
                                       (float)((q.y - L.NA.cy) * a.sw), zw.x, zw.y)
                         : make_float4((float)(q.x - L.NA.cx), (float)(q.y - L.NA.cy), zw.x, zw.y);
      wave_sync();
      if (t + 1 < tb) load(t + 1);
      if (L.tact) {
        if (ser)
          lds_sum32<true, UNI>(L.lx, L.ly, sf + L.rep * L.stride, tc, L.rep, L.nrep, ax, ay);
        else
          lds_sum32<false, UNI>(L.lx, L.ly, sf + L.rep * L.stride, tc, L.rep, L.nrep, ax, ay);
      }
    } else {
      double4 *sd = reinterpret_cast<double4 *>(st);
      if (lane < tc) sd[L.spos] = q;
      wave_sync();
      if (t + 1 < tb) load(t + 1);
      if (L.tact) lds_sum64(L.tx, L.ty, sd + L.rep * L.stride, tc, L.rep, L.nrep, ax, ay);
    }
    wave_sync();
  }
}

// A node C all of whose descendants fail the far test is summed as one run
// of sources.  For a descendant C' of C, |c' - c_A| <= d + R_C - R_C' < d + R_C
// =: reach, so
//   d' - R_A - R_C' < r_cut      when reach - R_A <= r_cut, and
//   R_A + R_C' > theta d'         when theta reach <= R_A,
// i.e. when d <= max(r_cut + R_A - R_C, R_A / theta - R_C) (compared squared).
__device__ inline bool whole_near(double inv_theta, double RA, double rcut, const NodeRec &NA,
                                  const NodeRec &C) {
  const double dx = C.cx - NA.cx, dy = C.cy - NA.cy;
  const double RC = C.region * 0.70710678118654752;
  const double m = fmax(rcut + RA - RC, RA * inv_theta - RC);
  return m >= 0.0 && dx * dx + dy * dy <= m * m;
}

// Near items are leaves (one step), whole near subtrees (one step per tile
// of 64 sources), smooth cells (one step; summed in batches, see
// smooth_batch), smooth runs: a whole near subtree whose blobs all lie in
// smooth-level nodes, one step per tile of 64 consecutive smooth-level nodes
// (smooth cells and small leaves), and leaf runs: near leaves that follow
// each other among the children of one node (only empty children between
// them), whose sources are one contiguous run -- summed like a whole near
// subtree, one step per tile of 64 (the record passed for a leaf run holds
// the first blob and the total count).
enum ItemKind { kItemLeaf = 0, kItemRange = 1, kItemSmooth = 2, kItemSmoothRun = 3, kItemRun = 4 };
__device__ inline bool tiled_run(int kind) { return kind == kItemRange || kind == kItemRun; }
__device__ inline int item_tiles(const FmmArgs &a, int c, const NodeRec &C, int kind) {
  if (kind == kItemSmoothRun) {
    const int2 r = ldro(a.srange + c);
    return (rfl(r.y) - rfl(r.x) + 63) >> 6;
  }
  return tiled_run(kind) ? (node_count(C.cnt) + 63) >> 6 : 1;
}

// ---- smooth cells ----------------------------------------------------------
// A near node of the smooth level (region in (smooth_region/2, smooth_region])
// holding at least smooth_min blobs is summed for all targets through the
// Taylor coefficients of smooth_coef_kernel (expansion_kernels.hip).  Only
// the mixed form with one core radius builds them (smooth_region 0 otherwise).
__device__ inline bool smooth_cell(const FmmArgs &a, const NodeRec &C) {
  return a.smooth_region > 0.0 && C.region <= a.smooth_region &&
         2.0 * C.region > a.smooth_region && node_count(C.cnt) >= a.smooth_min;
}

// Coefficients of phi^(q)(w) = (-1)^q sum_k (-w)^k / (k! (k + q + 1)) below
// kSmoothW0 (14 terms: the next is below 2e-10 of the sum there).
constexpr float kSmoothW0 = 1.5f;
constexpr int kSmoothSeries = 14;
struct SmoothSeries {
  float v[kSmoothSeries];
};
constexpr SmoothSeries make_smooth_series() {
  SmoothSeries t{};
  double f = 1.0;
  for (int k = 0; k < kSmoothSeries; k++) {
    if (k > 0) f *= (double)k;
    const double c = 1.0 / (f * (double)(k + kSmoothQ + 1));
    t.v[k] = (float)((k & 1) ? -c : c);
  }
  return t;
}
constexpr SmoothSeries kSmoothSer = make_smooth_series();

// One smooth cell at one target, FP32: Z' = (zx, zy) is the target relative
// to the cell centre in the scaled coordinates (sigma / sqrt(log2 e) units),
// row[] the cell's coefficients.  With w = ln2 |Z'|^2 and E = e^-w = 2^-|Z'|^2,
//   F = zb sum_m phi^(m)(w) nu_m zb^m + E sum_b y^b sum_a d_ab x^a    (Horner),
// phi^(m)(w) = (-1)^m int_0^1 t^m e^{-wt} dt from w phi^(m) + m phi^(m-1) =
// (-1)^(m+1) E: upward from phi = (1 - E)/w for w >= kSmoothW0, downward from
// the series of phi^(q) below it (each direction only where it is stable).
// Adds (Im F, Re F) to (gx, gy).
__device__ inline void smooth_eval(float zx, float zy, const float4 *row, float &gx, float &gy) {
  constexpr int Q = kSmoothQ;
  auto co = [&](int k) -> float2 {
    const float4 v = row[k >> 1];
    return (k & 1) ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
  };
  const float wp = fmaf(zx, zx, zy * zy);
  const float E = __builtin_amdgcn_exp2f(-wp);
  const float w = wp * 0.693147180559945309f;
  float ph[Q + 1];
  if (w >= kSmoothW0) {
    const float rw = __builtin_amdgcn_rcpf(w);
    ph[0] = (1.0f - E) * rw;
#pragma unroll
    for (int m = 1; m <= Q; m++) ph[m] = fmaf(-(float)m, ph[m - 1], (m & 1) ? E : -E) * rw;
  } else {
    float sr = kSmoothSer.v[kSmoothSeries - 1];
#pragma unroll
    for (int k = kSmoothSeries - 2; k >= 0; k--) sr = fmaf(sr, w, kSmoothSer.v[k]);
    ph[Q] = (Q & 1) ? -sr : sr;
#pragma unroll
    for (int m = Q; m >= 1; m--) ph[m - 1] = fmaf(-w, ph[m], (m & 1) ? E : -E) * (1.0f / (float)m);
  }
  // n = 0 column: acc = sum_m nu_m phi^(m) zb^m, S0 = acc zb
  float2 c = co(0);
  float ar = c.x * ph[Q], ai = c.y * ph[Q];
#pragma unroll
  for (int m = Q - 1; m >= 0; m--) {
    c = co(Q - m);
    float tr = fmaf(ar, zx, fmaf(ai, zy, c.x * ph[m]));
    float ti = fmaf(ai, zx, fmaf(-ar, zy, c.y * ph[m]));
    if ((Q - m) & 1) asm volatile("" : "+v"(tr), "+v"(ti) : : "memory");
    ar = tr;
    ai = ti;
  }
  float s0r = fmaf(ar, zx, ai * zy), s0i = fmaf(ai, zx, -ar * zy);
  asm volatile("" : "+v"(s0r), "+v"(s0i) : : "memory");
  // n >= 1: P = sum_b y^b R_b(x), R_b(x) = sum_a d_ab x^a (real monomials:
  // each Horner step multiplies a complex value by a real coordinate)
  float pr = 0.f, pi = 0.f;
  int k = Q + 1;
#pragma unroll
  for (int b = Q - 1; b >= 0; b--) {
    c = co(k++);
    float qr = c.x, qi = c.y;
#pragma unroll
    for (int i = Q - 2 - b; i >= 0; i--) {
      c = co(k++);
      qr = fmaf(qr, zx, c.x);
      qi = fmaf(qi, zx, c.y);
    }
    pr = fmaf(pr, zy, qr);
    pi = fmaf(pi, zy, qi);
    // the reads of the rows after the next stay after this row's arithmetic
    // (all rows' reads at once would hold ~90 registers): an empty asm
    // statement every second row that takes the running sum and orders memory
    if (b & 1) asm volatile("" : "+v"(pr), "+v"(pi) : : "memory");
  }
  gx += fmaf(E, pi, s0i);
  gy += fmaf(E, pr, s0r);
}

// Work units of one smooth evaluation: kSmoothUnits (about its cost in pair
// steps).
//
// Stage layout of the mixed form (float4 slots of one wave's slice):
//   [0, 128)    near-step sources while a near item is summed; the values of
//               one smooth batch (float2 per (target, cell) evaluation, at
//               most 3 x 64) while a batch is summed -- never both;
//   [128, 266)  coefficients of the pending smooth cells, kSmoothF4 per cell,
//               written by direct global-to-LDS loads when a cell is found;
//   [266, 268)  their offsets from the target leaf (float2 per cell);
//   [268, 300)  the targets' coordinates (float2 per target).
constexpr int kSmoothCoef = 128;
constexpr int kSmoothOff = 266;
constexpr int kSmoothTl = 268;
static_assert(kStageSlots <= kSmoothCoef && 2 * kSmoothCoef >= kSmoothBatch * 64,
              "near stage or smooth values overlap the coefficients");
static_assert(kSmoothCoef + kSmoothBatch * kSmoothF4 <= kSmoothOff &&
                  2 * kSmoothOff + kSmoothBatch <= 2 * kSmoothTl &&
                  kSmoothTl + 32 <= kStageSlots + kSmoothResSlots,
              "smooth stage exceeds the slice");
// Node k (lane k) of tile t of smooth run c: its record (cnt 0 past the end).
__device__ inline NodeRec srun_node(const FmmArgs &a, int c, int t, int lane, int &idx) {
  const int2 r = ldro(a.srange + c);
  idx = rfl(r.x) + 64 * t + lane;
  NodeRec C = {};
  if (idx < rfl(r.y)) C = a.nodes[idx];
  return C;
}

// Smooth cell C (node index cell) becomes cell k of the pending batch: lanes
// 0..kSmoothF4-1 load its coefficients straight into the stage (global to
// LDS, no registers; they land while the walk goes on), lane 0 writes its
// offset from the target leaf (FP64 difference, then FP32 scaled).
__device__ inline void smooth_fetch(const FmmArgs &a, const LeafCtx &L, int cell, const NodeRec &C,
                                    int k, int lane, float4 *st) {
  if (lane == 0)
    reinterpret_cast<float2 *>(st + kSmoothOff)[k] =
        make_float2((float)((C.cx - L.NA.cx) * a.sw), (float)((C.cy - L.NA.cy) * a.sw));
  if (lane < kSmoothF4)
    __builtin_amdgcn_global_load_lds(
        (const void *)(a.scoef + (size_t)cell * kSmoothF4 + lane),
        (__attribute__((address_space(3))) void *)(st + kSmoothCoef + k * kSmoothF4), 16, 0, 0);
}

// The nb pending cells (coefficients and offsets in the stage, see
// smooth_fetch): the nb x cA evaluations e = (target e mod cA, cell e / cA)
// are dealt to the lanes in rounds (lane e starts at its own (ti, rep)); each
// value goes to the stage's value slots, and the lane of target t (replica 0)
// adds its nb values in cell order to its FP64 sums -- a cell adds the same
// whatever batch it falls in.
__device__ inline void smooth_batch(const LeafCtx &L, int nb, int lane, float4 *st, double &ax,
                                    double &ay) {
  float2 *res = reinterpret_cast<float2 *>(st);
  const float4 *coef = st + kSmoothCoef;
  const float2 *off = reinterpret_cast<const float2 *>(st + kSmoothOff);
  float2 *tl = reinterpret_cast<float2 *>(st + kSmoothTl);
  if (L.rep == 0) tl[L.ti] = make_float2(L.lx, L.ly);
  // the direct loads of the coefficients have landed (vmcnt 0), then the
  // wave's LDS writes are visible to all its lanes
  __builtin_amdgcn_s_waitcnt(0x0F70);
  wave_sync();
  const int cA = L.cA, ne = nb * cA;
  const int q64 = L.nrep, r64 = 64 - q64 * cA;  // 64 / cA, 64 % cA (cA <= 64)
  int t = L.ti, k = L.rep;
  for (int e = lane; e < ne; e += 64) {
    const float4 *row = coef + k * kSmoothF4;
    const float2 o = off[k];
    const float2 p = tl[t];
    float gx = 0.f, gy = 0.f;
    smooth_eval(p.x - o.x, p.y - o.y, row, gx, gy);
    res[e] = make_float2(gx, gy);
    t += r64;
    k += q64;
    if (t >= cA) {
      t -= cA;
      k++;
    }
  }
  wave_sync();
  if (L.tact && L.rep == 0) {
    for (int kk = 0; kk < nb; kk++) {
      const float2 g = res[kk * cA + L.ti];
      ax += (double)g.x;
      ay += (double)g.y;
    }
  }
  wave_sync();
}

// Steps ta..tb-1 of near item c.
template <bool MIXED, bool UNI>
__device__ inline void near_steps(const FmmArgs &a, const LeafCtx &L, int c, const NodeRec &C,
                                  int kind, bool ser, int ta, int tb, int lane,
                                  typename StageOf<MIXED>::T *st, double &ax,
                                  double &ay) {
  if (tiled_run(kind)) {
    const int s0 = kind == kItemRun ? rfl(C.first) : rfl(ldro(a.nfirst + c));
    near_range<MIXED, UNI>(a, L, s0, node_count(C.cnt), ser, ta, tb, lane, st, ax, ay);
  } else {
    near_leaf<MIXED, UNI>(a, L, C, ser, lane, st, ax, ay);
  }
}

// Steps ta..tb-1 of near item c of any kind; smooth cells go to push(cell)
// (the caller's batch), the small leaves of a smooth run are summed here.
template <bool MIXED, bool UNI, typename Push>
__device__ inline void item_steps(const FmmArgs &a, const LeafCtx &L, int c, const NodeRec &C,
                                  int kind, bool ser, int ta, int tb, int lane,
                                  typename StageOf<MIXED>::T *st, double &ax, double &ay,
                                  Push &&push) {
  if constexpr (MIXED && UNI) {
    if (kind == kItemSmooth) {
      push(c, C);
      return;
    }
    if (kind == kItemSmoothRun) {
      for (int t = ta; t < tb; t++) {
        int idx;
        const NodeRec N = srun_node(a, c, t, lane, idx);
        const int cnt = node_count(N.cnt);
        uint64_t m = __builtin_amdgcn_ballot_w64(cnt > 0);
        const uint64_t sm = __builtin_amdgcn_ballot_w64(cnt >= a.smooth_min);
        // the series test of each small leaf, one per lane
        const uint64_t serm =
            __builtin_amdgcn_ballot_w64(cnt > 0 && cnt < a.smooth_min && needs_series(a, L, N));
        while (m) {
          const int k = __builtin_ctzll(m);
          m &= m - 1;
          const int node = lane_of(idx, k);
          if ((sm >> k) & 1) {
            push(node, node_of_lane(N, k));
          } else {
            near_leaf<MIXED, UNI>(a, L, node_of_lane(N, k), (serm >> k) & 1, lane, st, ax, ay);
          }
        }
      }
      return;
    }
  }
  near_steps<MIXED, UNI>(a, L, c, C, kind, ser, ta, tb, lane, st, ax, ay);
}


// The far test of a cell C for target leaf A: R_A + R_C <= theta d and
// d - R_A - R_C >= r_cut, i.e. d >= max((R_A + R_C) / theta, R_A + R_C + r_cut),
// compared squared (no square root).  Every walk -- fused pass, chunks,
// counting -- calls this one function, so they all make the same decisions.
__device__ inline bool far_cell(double inv_theta, double RA, double rcut, const NodeRec &NA,
                                const NodeRec &C) {
  const double dx = C.cx - NA.cx, dy = C.cy - NA.cy;
  const double sr = RA + C.region * 0.70710678118654752;
  const double lim = fmax(sr * inv_theta, sr + rcut);
  return dx * dx + dy * dy >= lim * lim;
}

// The walk of one target leaf, level-parallel: `stack` (this wave's LDS,
// kStack entries) holds the first-child index of every open node still to be
// expanded.  A step takes up to 16 of them (the latest pushed) and classifies
// their 64 children at once, lane = child: far cells go to far(mask, c) (the
// lanes of `mask` hold one each) and near items to item(c, C, kind) in lane
// order; open children are pushed.
// item returning false ends the walk.  The order depends on the tree and the
// target leaf only, so every pass over a leaf (fused, continuation, inline)
// meets the same items in the same order.  Pops shrink to one node per step
// near the top of the stack; a depth-first descent then adds at most 3
// entries per level, which kStackMargin covers.
constexpr int kStack = 384;
constexpr int kStackMargin = 96;
constexpr bool kMergeLeaves = true;  // consecutive sibling near leaves as one run

// ---- local expansions inherited from the parent (L2L) ----------------------
// A far cell of target leaf A that is also far from A's parent P is summed
// once into P's local expansion (parent_local_kernel) and reaches A, and every
// other target-leaf child of P, through the exact shift L2L.  To count each
// cell exactly once the parent takes only cells that every one of its
// target-leaf children would have taken as far itself: P opens a node iff all
// of them open it (the walk's own rule), and lists an opened node's child C iff
// it is far from P with rcut_P = the largest near radius of those children.
// Far from P implies far from each child (a child's centre is R_P / 2 from
// P's, R_A = R_P / 2 and theta <= 1), so every child, whose walk meets C below
// a node all of them opened, would have listed it: its walk skips such a C
// ("covered") instead.  Stack entries carry the flag "the parent opened this
// node" in bit 31.  Only parents with at least two target-leaf children take
// part.
//
// The parent's walk also lists where its children's walks start (pfront):
// every group of four siblings below a node all of them open that holds a
// child neither covered nor opened by all of them (a "boundary" child), with
// a mask of the children to skip (covered or opened by all; each one's part
// is the parent's, or lies in later entries).  A child's walk then starts
// from these entries, not from the root: it meets exactly the nodes the
// walk from the root meets below the nodes all children open, classified by
// the same rule, and no covered cell.  Entries: first-child index (below
// 2^27) | skip mask << 27.
constexpr int kCovBit = int(0x80000000u);
constexpr int kSkipShift = kSkipBits;
constexpr int kIndexMask = (1 << kSkipShift) - 1;
// The parent's box and the target-leaf siblings of A, kept in the wave's LDS
// (read only where a walk step meets a node the parent opened).
struct ParentBox {
  double npx, npy, nregion, RP, rcutP;
  double sx[4], sy[4], srcut[4];
};
struct ParentCtx {
  bool on = false;
  int g = 0;               // the sibling group of A (the parent's expansion slot)
  int pP = 0;              // terms of the parent's expansion
  uint32_t sib = 0;        // A's other target-leaf siblings (mask over q)
  const ParentBox *box = nullptr;
  const uint32_t *front = nullptr;  // where the walk starts (null: the root)
  int nfront = 0;
};

// The start list of target leaf li's walk, or null (walk from the root).
__device__ inline const uint32_t *leaf_front(const FmmArgs &a, int li, int &nfront) {
  nfront = 0;
  if (!a.l2l || !a.pfront) return nullptr;
  const int g = rfl(ldro(a.leaves + li)) >> 2;
  const int n = rfl(ldro(a.pfcnt + g));
  if (n < 0) return nullptr;
  nfront = n;
  return a.pfront + (size_t)g * kFrontCap;
}

// Near radius and terms of target leaf li.
__device__ inline void leaf_rcut_p(const FmmArgs &a, int li, double &rcut, int &p) {
  rcut = a.rcut;
  p = a.p;
  if (a.leaf_par) {
    const float2 lp = a.leaf_par[li];
    p = min(a.p, max(2, (int)lp.x));
    rcut = (double)lp.y * a.smax;
  }
}

// The target-leaf children of sibling group g (mask over q), their near radii
// and terms (only those of the mask are set).
__device__ inline uint32_t group_targets(const FmmArgs &a, int g, NodeRec C[4], double rc[4],
                                         int pq[4]) {
  uint32_t m = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    C[q] = ldro(a.nodes + 4 * g + q);
    rc[q] = 0.0;
    pq[q] = 0;
    if (!node_internal(C[q].cnt) && node_count(C[q].cnt) > 0) {
      m |= 1u << q;
      leaf_rcut_p(a, rfl(ldro(a.nodeli + 4 * g + q)), rc[q], pq[q]);
    }
  }
  return m;
}

__device__ inline ParentCtx parent_ctx(const FmmArgs &a, int li, int lane, ParentBox *box) {
  ParentCtx pc;
  if (!a.l2l) return pc;
  const int iA = rfl(ldro(a.leaves + li));
  const int g = iA >> 2;
  const int par = rfl(ldro(a.gparent + g));
  if (par < 0) return pc;
  NodeRec C[4];
  double rc[4];
  int pq[4];
  const uint32_t m = group_targets(a, g, C, rc, pq);
  if (__builtin_popcount(m) < a.l2l) return pc;
  pc.on = true;
  pc.g = g;
  pc.sib = m & ~(1u << (iA & 3));
  pc.box = box;
  if (a.pfront) {
    const int n = rfl(ldro(a.pfcnt + g));
    if (n >= 0) {
      pc.front = a.pfront + (size_t)g * kFrontCap;
      pc.nfront = n;
    }
  }
  const NodeRec NP = ldro(a.nodes + par);
  double rcutP = 0.0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if ((m >> q) & 1) {
      rcutP = fmax(rcutP, rc[q]);
      pc.pP = max(pc.pP, pq[q]);
    }
  }
  if (lane == 0) {
    box->npx = NP.cx;
    box->npy = NP.cy;
    box->nregion = NP.region;
    box->RP = NP.region * 0.70710678118654752;
    box->rcutP = rcutP;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      box->sx[q] = C[q].cx;
      box->sy[q] = C[q].cy;
      box->srcut[q] = rc[q];
    }
  }
  wave_sync();
  return pc;
}

__device__ inline NodeRec parent_node(const ParentCtx &pc) {
  NodeRec N = {};
  N.cx = pc.box->npx;
  N.cy = pc.box->npy;
  N.region = pc.box->nregion;
  return N;
}

// Whether a leaf box of half diagonal RA at (x, y) with near radius rcut opens
// internal node c (the walk's rule: not far, not a smooth cell, not a whole
// near subtree taken as one item).
__device__ inline bool leaf_opens(const FmmArgs &a, bool smooth_on, double inv_theta, double RA,
                                  double rcut, double x, double y, double region, int c,
                                  const NodeRec &C) {
  NodeRec B = {};
  B.cx = x;
  B.cy = y;
  B.region = region;
  if (far_cell(inv_theta, RA, rcut, B, C)) return false;
  if (smooth_on && smooth_cell(a, C)) return false;
  if (whole_near(inv_theta, RA, rcut, B, C) && (!smooth_on || a.srange[c].x != -2)) return false;
  return true;
}

// The walk of one target leaf, level-parallel: `stack` (this wave's LDS,
// kStack entries) holds the first-child index of every open node still to be
// expanded.  A step takes up to 16 of them (the latest pushed) and classifies
// their 64 children at once, lane = child: far cells go to far(mask, c) (the
// lanes of `mask` hold one each) and near items to item(c, C, kind) in lane
// order; open children are pushed.
// item returning false ends the walk.  The order depends on the tree and the
// target leaf only, so every pass over a leaf (fused, continuation, inline)
// meets the same items in the same order (cells covered by the parent are far
// cells, never items).  Pops shrink to one node per step near the top of the
// stack; a depth-first descent then adds at most 3 entries per level, which
// kStackMargin covers.  With L2L, children of nodes the parent opened that are
// far from the parent are skipped (the parent's expansion has them).
template <bool SMOOTH, bool L2L, typename FarFn, typename ItemFn>
__device__ inline void walk_wide(const FmmArgs &a, const LeafCtx &L, int lane, int *stack,
                                 FarFn &&far, ItemFn &&item, const ParentCtx &pc) {
  const bool smooth_on = SMOOTH && a.smooth_region > 0.0;
  // from the parent's start list the walk never meets a covered cell
  const bool cov_on = L2L && pc.on && !pc.front;
  int ns = 1;
  if (pc.front) {
    // reversed: the walk pops the top first, so it takes the groups in the
    // parent's walk order (measured: the continuation chunks' walks reach
    // their items sooner this way; 4M mixed form without smooth cells 35.4 ->
    // 31.5 ms per step)
    ns = pc.nfront;
    for (int i = lane; i < ns; i += 64) stack[i] = (int)pc.front[ns - 1 - i];
  } else if (lane == 0) {
    stack[0] = ldro(a.nodes + rfl(root_of(a, L.NA.first))).first | (cov_on ? kCovBit : 0);
  }
  wave_sync();
  while (ns > 0) {
    const int k = min(min(16, ns), max(1, (kStack - kStackMargin - ns) / 3));
    int c = -1;
    bool cov = false;
    if ((lane >> 2) < k) {
      const int e = stack[ns - k + (lane >> 2)];
      cov = e < 0;
      c = (e & kIndexMask) + (lane & 3);
      if ((e >> (kSkipShift + (lane & 3))) & 1) c = -1;  // skipped child of a start entry
    }
    wave_sync();
    ns -= k;
    NodeRec C = {};
    if (c >= 0) C = a.nodes[c];
    bool f = false, it = false, op = false;
    int kind = kItemLeaf;
    if (node_count(C.cnt) > 0) {
      if (cov_on && cov && far_cell(L.inv_theta, pc.box->RP, pc.box->rcutP, parent_node(pc), C)) {
        // covered by the parent's expansion
      } else if (far_cell(L.inv_theta, L.RA, L.rcut, L.NA, C)) {
        f = true;
      } else if (smooth_on && smooth_cell(a, C)) {
        it = true;
        kind = kItemSmooth;
      } else if (!node_internal(C.cnt)) {
        it = true;
      } else if (whole_near(L.inv_theta, L.RA, L.rcut, L.NA, C) &&
                 (!smooth_on || a.srange[c].x != -2)) {
        // whole near subtree: one run of sources, or of smooth-level nodes
        // if all its blobs lie in them; with smooth cells among other blobs
        // it is opened
        it = true;
        kind = (smooth_on && a.srange[c].x >= 0) ? kItemSmoothRun : kItemRange;
      } else {
        op = true;
      }
    }
    const uint64_t fm = __builtin_amdgcn_ballot_w64(f);
    const uint64_t im = __builtin_amdgcn_ballot_w64(it);
    const uint64_t om = __builtin_amdgcn_ballot_w64(op);
    // the series test of each near leaf or range (mixed form), one per lane
    const uint64_t sm = __builtin_amdgcn_ballot_w64(
        it && (kind == kItemLeaf || kind == kItemRange) && needs_series(a, L, C));
    if (op) {
      int e = C.first;
      if (cov_on && cov) {
        // the parent opens C iff every target-leaf child does
        bool all = true;
#pragma unroll
        for (int q = 0; q < 4; q++)
          if ((pc.sib >> q) & 1)
            all = all && leaf_opens(a, smooth_on, L.inv_theta, L.RA, pc.box->srcut[q],
                                    pc.box->sx[q], pc.box->sy[q], L.NA.region, c, C);
        if (all) e |= kCovBit;
      }
      const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(om >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)om, 0));
      if (ns + below < kStack) stack[ns + below] = e;
    }
    ns += __builtin_popcountll(om);
    wave_sync();
    if (fm) far(fm, c);
    // near leaves in one group of four children (lanes 4j..4j+3) with only
    // empty children between them are one item (kItemRun)
    const uint64_t lm = kMergeLeaves ? __builtin_amdgcn_ballot_w64(it && kind == kItemLeaf) : 0;
    const uint64_t em =
        kMergeLeaves ? __builtin_amdgcn_ballot_w64(c >= 0 && node_count(C.cnt) == 0) : 0;
    uint64_t m = im;
    while (m) {
      const int b = __builtin_ctzll(m);
      m &= m - 1;
      NodeRec Cb = node_of_lane(C, b);
      int kb = lane_of(kind, b);
      bool ser = (sm >> b) & 1;
      if (kMergeLeaves && kb == kItemLeaf) {
        int tot = node_count(Cb.cnt);
        for (int b2 = b + 1; b2 <= (b | 3); b2++) {
          if ((lm >> b2) & 1) {
            tot += node_count(lane_of(C.cnt, b2));
            ser = ser || ((sm >> b2) & 1);
            m &= ~(1ull << b2);
            kb = kItemRun;
          } else if (!((em >> b2) & 1)) {
            break;
          }
        }
        if (kb == kItemRun) Cb.cnt = tot;
      }
      if (!item(lane_of(c, b), Cb, kb, ser)) return;
    }
  }
}

// A's local expansion at the start of its walk: with L2L, the parent's
// (scaled by s_P about c_P) shifted to A's centre,
//   Lt_A[m] = 2^-m sum_{l=m}^{pP-1} Lt_P[l] C(l, m) t^(l-m),  t = (c_A - c_P) / s_P
// (s_A = s_P / 2; t = (+-1/4, +-1/4) exactly), else zero.
__device__ inline void l2l_init(const FmmArgs &a, const ParentCtx &pc, const LeafCtx &L, int lane,
                                double2 *sloc) {
  if (lane < kFmmMaxP) {
    double2 v = make_double2(0.0, 0.0);
    if (pc.on && lane < pc.pP) {
      const double2 *lp = a.plocal + (size_t)pc.g * kFmmMaxP;
      const ParentBox *b = pc.box;
      const double tr = (L.NA.cx - b->npx) / b->nregion, ti = (L.NA.cy - b->npy) / b->nregion;
      const int m = lane;
      double sr = 0.0, si = 0.0;
      for (int l = pc.pP - 1; l >= m; l--) {
        const double2 q = lp[l];
        const double b = kM2lBinom.v[m][l - m];  // C(l, m)
        const double nr = sr * tr - si * ti + b * q.x;
        const double ni = sr * ti + si * tr + b * q.y;
        sr = nr;
        si = ni;
      }
      const double sc = ldexp(1.0, -m);
      v = make_double2(sr * sc, si * sc);
    }
    sloc[lane] = v;
  }
  wave_sync();
}

// Work units of tiles [0, t) of item c (nt tiles): near_units per tile; a
// smooth run sums fixed per-node units (a smooth cell kSmoothUnits, a small
// leaf its blobs + kStepUnits) kept as prefix sums over the nodes (sunits).
__device__ inline int64_t item_prefix(const FmmArgs &a, const LeafCtx &L, int c, const NodeRec &C,
                                      int kind, int nt, int t) {
  if (t <= 0) return 0;
  if (kind == kItemSmooth) return L.su;
  if (kind == kItemLeaf) return near_units(node_count(C.cnt), L);
  if (tiled_run(kind)) {
    int64_t u = (int64_t)min(t, nt - 1) * near_units(64, L);
    if (t >= nt) u += near_units(node_count(C.cnt) - 64 * (nt - 1), L);
    return u;
  }
  const int2 r = ldro(a.srange + c);
  const int lo = rfl(r.x), hi = rfl(r.y);
  return (int64_t)rfl((int)(ldro(a.sunits + min(lo + 64 * t, hi)) - ldro(a.sunits + lo)));
}

// First tile t in [t0, nt) whose work before it, base + prefix(t), reaches v
// (nt if none); the prefix does not decrease in t.
__device__ inline int item_find(const FmmArgs &a, const LeafCtx &L, int c, const NodeRec &C,
                                int kind, int nt, int64_t base, int64_t v, int t0) {
  int lo = t0, hi = nt;  // answer in [lo, hi]
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (base + item_prefix(a, L, c, C, kind, nt, mid) >= v)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Sum over the replicas of each target; the result is valid in the lanes
// with rep == 0.  red = this wave's LDS slice of 64 entries.
__device__ inline double2 reduce_replicas(const LeafCtx &L, int lane, double2 *red, double ax,
                                          double ay) {
  wave_sync();
  red[lane] = make_double2(ax, ay);
  wave_sync();
  double sx = 0.0, sy = 0.0;
  if (L.tact && L.rep == 0) {
    for (int rr = 0; rr < L.nrep; rr++) {
      const double2 v = red[rr * L.cA + L.ti];
      sx += v.x;
      sy += v.y;
    }
  }
  wave_sync();
  return make_double2(sx, sy);
}

// Positions in a leaf's near work: (item ordinal in walk order) << 32 | tile.
__device__ inline int64_t work_pos(int k, int t) { return ((int64_t)k << 32) | (uint32_t)t; }

// Near work of one chunk: the tiles from position start up to (not
// including) end, met by walking the leaf again from the root (far cells are
// the fused pass's; items before the chunk are only counted); smooth cells in
// batches, the last batch summed at the end.
template <bool MIXED, bool UNI>
__device__ inline void run_chunk(const FmmArgs &a, const LeafCtx &L, int li, int64_t start,
                                 int64_t end, int lane, typename StageOf<MIXED>::T *st, int *stack,
                                 double &ax, double &ay) {
  constexpr bool SM = MIXED && UNI;
  // the fused pass's walk order: from the parent's start list if it had one
  ParentCtx pc;
  pc.front = leaf_front(a, li, pc.nfront);
  const int k0 = (int)(start >> 32), t0 = (int)(uint32_t)start;
  const int k1 = (int)(end >> 32), t1 = (int)(uint32_t)end;
  int k = 0;
  int snb = 0;
  const int sbsz = L.sbsz;
  auto push = [&](int cell, const NodeRec &C) {
    if constexpr (SM) {
      smooth_fetch(a, L, cell, C, snb, lane, st);
      if (++snb == sbsz) {
        smooth_batch(L, snb, lane, st, ax, ay);
        snb = 0;
      }
    }
  };
  walk_wide<SM, false>(a, L, lane, stack, [](uint64_t, int) {},
                [&](int c, const NodeRec &C, int kind, bool ser) -> bool {
    const int ki = k++;
    if (ki < k0) return true;
    if (ki > k1 || (ki == k1 && t1 == 0)) return false;
    const int nt = item_tiles(a, c, C, kind);
    const int ta = ki == k0 ? t0 : 0;
    const int tb = ki == k1 ? min(t1, nt) : nt;
    if (ta < tb) item_steps<MIXED, UNI>(a, L, c, C, kind, ser, ta, tb, lane, st, ax, ay, push);
    return ki < k1;
  }, pc);
  if constexpr (SM) {
    if (snb > 0) smooth_batch(L, snb, lane, st, ax, ay);
  }
}

// Far cells of a box X (a target leaf, or the parent whose local expansion
// its leaf children inherit), listed in LDS by a walk, summed in one burst
// into X's scaled local expansion sloc[l] (l < kFmmMaxP).  Per quad of cells,
// lane group mk (16 lanes, row i = mi) forms (s_X w)^(i+1) and
// (-1)^(i+1) a_i w^i of cell 4q + mk once; each MFMA step then takes the real
// and imaginary parts of two cells (K slots cell re, cell im, cell' re, cell'
// im) from the groups that formed them (lane permutes).  The next quad's cell
// data is loaded before this quad's arithmetic.  The burst adds its part of
// the local expansion, Lt_l = (1/s_X) sum_k C(l+k,k) G[l][k], to sloc.
// sXpow = s_X^(mi+1) of this lane.  Ends with a wave-level fence.
template <bool MIXED>
__device__ inline void far_contract(const FmmArgs &a, const NodeRec &NA, double sA, int pA,
                                   double2 sApow, int lane, const int *far_list, int nfar,
                                   double2 *sloc) {
  const int mi = lane & 15, mk = lane >> 4;
  const bool row_ok = mi < pA;
  FarAcc<MIXED> acc;
  const int nq = (nfar + 3) >> 2;
  // centre and coefficient of the quad's cell, loaded a quad ahead; a lane
  // without a cell (past the list, or row i >= p) loads node 0 and its
  // result is dropped, so the loads and the arithmetic take no branches
  double2 cc, ak;
  bool ok;
  auto fetch = [&](int q, double2 &c2, double2 &k, bool &v) {
    const int idx = 4 * q + mk;
    const int cell = idx < nfar ? far_list[idx] : -1;
    v = cell >= 0 && row_ok;
    const int cs = v ? cell : 0;
    c2 = make_double2(a.nodes[cs].cx, a.nodes[cs].cy);
    k = a.coef[(size_t)cs * a.p + (v ? mi : 0)];
  };
  fetch(0, cc, ak, ok);
  for (int q = 0; q < nq; q++) {
    double2 ncc = cc, nak = ak;
    bool nok = false;
    if (q + 1 < nq) fetch(q + 1, ncc, nak, nok);
    // w = 1 / (dr + i di); 1 / |d|^2 from the hardware reciprocal and two
    // Newton steps
    const double dr = cc.x - NA.cx, di = cc.y - NA.cy;
    const double r2 = fma(dr, dr, di * di);
    double inv = __builtin_amdgcn_rcp(r2);
    inv = inv * fma(-r2, inv, 2.0);
    inv = inv * fma(-r2, inv, 2.0);
    const double2 wv = make_double2(dr * inv, -di * inv);
    const double2 wi = cpow16(wv, mi);                          // w^i
    const double2 wi1 = cmul_fma(wi, wv);
    double2 pw = make_double2(wi1.x * sApow.x, wi1.y * sApow.x);  // (s_A w)^(i+1)
    double2 qv = cmul_fma(ak, wi);                                // a_i w^i
    if ((mi & 1) == 0) qv = make_double2(-qv.x, -qv.y);           // (-1)^(i+1)
    if (!ok) {
      pw = make_double2(0.0, 0.0);
      qv = make_double2(0.0, 0.0);
    }
    // operands rounded to the MFMA type before they cross lanes
    typedef typename FarAcc<MIXED>::T FT;
    const FT pwx = (FT)pw.x, pwy = (FT)pw.y, qvx = (FT)qv.x, qvy = (FT)qv.y;
    const bool im = (mk & 1) != 0;
    for (int st = 0; st < 2; st++) {
      if (4 * q + 2 * st >= nfar) continue;  // wave-uniform
      const int src = mi + 16 * (2 * st + (mk >> 1));
      const FT px = __shfl(pwx, src, 64), py = __shfl(pwy, src, 64);
      const FT qx = __shfl(qvx, src, 64), qy = __shfl(qvy, src, 64);
      acc.add(im ? py : px, im ? -qy : qx, im ? qx : qy);
      acc.tick();
    }
    cc = ncc;
    ak = nak;
    ok = nok;
  }
  acc.fold();
  // lane holds rows acc.row(mk, r), column mi of G (real part in r0,
  // imaginary in r1)
  for (int r = 0; r < 4; r++) {
    const int row = acc.row(mk, r);
    const double bc = kM2lBinom.v[row][mi];  // C(row + mi, mi)
    double gr = (mi < pA) ? acc.r0(r) * bc : 0.0;
    double gi = (mi < pA) ? acc.r1(r) * bc : 0.0;
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
      gr += __shfl_xor(gr, o, 64);
      gi += __shfl_xor(gi, o, 64);
    }
    if (mi == 0 && row < kFmmMaxP) {
      double2 v = sloc[row];
      v.x += gr / sA;
      v.y += gi / sA;
      sloc[row] = v;
    }
  }
  wave_sync();
}

// The fused pass for target leaf li (index in the leaf list), one wave.
template <bool MIXED, bool UNI>
__device__ __attribute__((always_inline)) void fmm_process_leaf(const FmmArgs &a, int li, int w, int lane,
                                 double2 (*s_loc)[kFmmMaxP], double2 (*s_red)[64],
                                 typename StageOf<MIXED>::T (*s_stage)[stage_slots<MIXED>()],
                                 int64_t (*s_ck)[kMaxCk], int32_t (*s_cku)[kMaxCk],
                                 int (*s_stack)[kStack], int (*s_far)[kFarList],
                                 ParentBox *s_pbox) {
  li = rfl(li);
  const LeafCtx L = leaf_ctx<MIXED, UNI>(a, li, lane);
  const NodeRec &NA = L.NA;
  const double sA = L.sA;
  const int pA = L.pA;
  double ax = 0.0, ay = 0.0;
  // s_A^(i+1) of this lane's MFMA row i = lane mod 16 (real)
  const double2 sApow = cpow_u(make_double2(sA, 0.0), (lane & 15) + 1);

  // The local expansion starts from the parent's, shifted to A's centre, when
  // the parent's far cells are summed once for its target-leaf children
  // (l2l_init); far cells of A's own are listed in LDS (s_far) by the walk and
  // summed in bursts (far_contract), at the end or when the list is full, so
  // the MFMA accumulators live only there.
  const ParentCtx pc = parent_ctx(a, li, lane, &s_pbox[w]);
  l2l_init(a, pc, L, lane, s_loc[w]);
  int nfar = 0;
  auto far_flush = [&]() {
    far_contract<MIXED>(a, NA, sA, pA, sApow, lane, s_far[w], nfar, s_loc[w]);
    nfar = 0;
  };
  auto far = [&](uint64_t fm, int c) {
    const int nf = __builtin_popcountll(fm);
    if (nfar + nf > a.far_burst) far_flush();
    if ((fm >> lane) & 1)
      s_far[w][nfar + __builtin_amdgcn_mbcnt_hi((uint32_t)(fm >> 32),
                                                __builtin_amdgcn_mbcnt_lo((uint32_t)fm, 0))] = c;
    nfar += nf;
    wave_sync();
  };

  // near work: summed here up to the first threshold, then only counted,
  // with a chunk boundary every `step` units (kept in LDS; when the slots run
  // out, every second boundary is dropped and the step doubles)
  const bool split = a.split_budget > 0;
  // the mixed form's chunks are an eighth of the first budget: few leaves
  // are split there, and the continuation pass then ends with its longest
  // chunk (walking the leaf again up to it, then summing it, on one wave)
  int64_t cum = 0, step = split ? (MIXED ? max(a.split_budget / 8, 1) : a.split_budget) : 0;
  int64_t thr = split ? (int64_t)a.split_budget : INT64_MAX;
  int nck = 0;
  constexpr bool SM = MIXED && UNI;
  // pending smooth cells (lane k holds cell k of the batch); only the summing
  // part of the walk adds cells, so the last batch is summed after the walk
  int snb = 0;
  const int sbsz = L.sbsz;
  auto push = [&](int cell, const NodeRec &C) {
    if constexpr (SM) {
      smooth_fetch(a, L, cell, C, snb, lane, s_stage[w]);
      if (++snb == sbsz) {
        smooth_batch(L, snb, lane, s_stage[w], ax, ay);
        snb = 0;
      }
    }
  };
  int nitem = 0;
  auto item = [&](int c, const NodeRec &C, int kind, bool ser) -> bool {
    const int ki = nitem++;
    const int nt = item_tiles(a, c, C, kind);
    const int64_t c0 = cum;
    const int64_t total = item_prefix(a, L, c, C, kind, nt, nt);
    int t = 0;
    if (nck == 0) {
      // the tiles that start below the first threshold are summed here
      t = c0 + total <= thr ? nt : item_find(a, L, c, C, kind, nt, c0, thr, 0);
      if (t > 0)
        item_steps<MIXED, UNI>(a, L, c, C, kind, ser, 0, t, lane, s_stage[w], ax, ay, push);
    }
    // a chunk boundary at every tile whose work before it reaches thr
    while (t < nt && c0 + total > thr) {
      t = item_find(a, L, c, C, kind, nt, c0, thr, t);
      if (t >= nt) break;
      if (nck == kMaxCk) {
        int64_t keep = 0;
        int32_t keepu = 0;
        if (lane < kMaxCk / 2) {
          keep = s_ck[w][2 * lane];
          keepu = s_cku[w][2 * lane];
        }
        wave_sync();
        if (lane < kMaxCk / 2) {
          s_ck[w][lane] = keep;
          s_cku[w][lane] = keepu;
        }
        wave_sync();
        nck = kMaxCk / 2;
        step *= 2;
        thr = (int64_t)rfl64((uint64_t)s_cku[w][nck - 1]) + step;
        continue;
      }
      // a boundary: its position (where the chunk starts) and its work count
      const int64_t at = c0 + item_prefix(a, L, c, C, kind, nt, t);
      if (lane == 0) {
        s_ck[w][nck] = work_pos(ki, t);
        s_cku[w][nck] = (int32_t)min(at, (int64_t)INT32_MAX);  // saturated: only a step base
      }
      nck++;
      thr = at + step;
    }
    cum = c0 + total;
    return true;
  };
  int *stack = s_stack[w];
  walk_wide<SM, true>(a, L, lane, stack, far, item, pc);
  if constexpr (SM) {
    if (snb > 0) smooth_batch(L, snb, lane, s_stage[w], ax, ay);
  }
  if (nfar > 0) far_flush();

  // s_loc[w], s_red[w], s_stage[w] and s_ck[w] belong to this wave alone: a
  // wave-level fence orders the LDS writes of some lanes before the reads of
  // others
  wave_sync();
  double phr = 0.0, phi = 0.0;
  if (L.tact && L.rep == 0) {
    // L2P: sum_l Lt_l u^l, u = (z - c_A)/s_A
    // the target's index formed anew (opaque to the compiler), so that no
    // copy of it is kept through the walk
    int tj = L.tj;
    asm volatile("" : "+v"(tj));
    const double tx = MIXED ? a.sx[tj] : L.tx, ty = MIXED ? a.sy[tj] : L.ty;
    const double ur = (tx - NA.cx) / sA, ui = (ty - NA.cy) / sA;
    double2 s = s_loc[w][pA - 1];
    for (int l = pA - 2; l >= 0; l--) {
      const double2 Lc = s_loc[w][l];
      s = make_double2(s.x * ur - s.y * ui + Lc.x, s.x * ui + s.y * ur + Lc.y);
    }
    phr = s.x;
    phi = s.y;
  }
  const double2 S = reduce_replicas(L, lane, s_red[w], ax, ay);

  // hand the chunks (positions s_ck[k] to s_ck[k + 1]) to the continuation
  // pass; without room for them, run them here with the same arithmetic
  bool inline_chunks = false;
  double2 extra = make_double2(0.0, 0.0);
  if (split) {
    int2 rec = make_int2(0, 0);
    if (nck > 0) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(a.ck_count, (uint32_t)nck);
      base = (uint32_t)rfl((int)__shfl((int)base, 0, 64));
      const bool fits = (int64_t)base + nck <= (int64_t)a.ck_cap;
      for (int k = lane; k < nck; k += 64) {
        if ((int64_t)base + k < (int64_t)a.ck_cap) {
          SplitCk r;
          r.cum = s_ck[w][k];
          r.end = k + 1 < nck ? s_ck[w][k + 1] : INT64_MAX;
          r.li = fits ? li : -1;  // the continuation pass skips li < 0
          r.pad = 0;
          a.ck[base + k] = r;
        }
      }
      if (fits) {
        rec = make_int2(a.ck_base + (int)base, nck);
      } else {
        inline_chunks = true;
        for (int k = 0; k < nck; k++) {
          const int64_t start = (int64_t)rfl64((uint64_t)s_ck[w][k]);
          const int64_t end = k + 1 < nck ? (int64_t)rfl64((uint64_t)s_ck[w][k + 1]) : INT64_MAX;
          double cx = 0.0, cy = 0.0;
          run_chunk<MIXED, UNI>(a, L, li, start, end, lane, s_stage[w], stack, cx, cy);
          const double2 P = reduce_replicas(L, lane, s_red[w], cx, cy);
          extra.x += P.x;
          extra.y += P.y;
        }
      }
    }
    if (lane == 0) {
      a.leaf_ck[li] = rec;
      // the leaves with records, for the chunk sums (their order there does
      // not matter: each leaf's sums are added in its own chunk order)
      if (rec.y > 0) a.split_list[atomicAdd(a.split_count, 1u)] = li;
    }
  }
  if (L.tact && L.rep == 0) {
    double u = 0.5 * phi / M_PI + 0.5 * S.x / M_PI;
    double v = 0.5 * phr / M_PI + 0.5 * S.y / M_PI;
    if (inline_chunks) {
      u = u + 0.5 * extra.x / M_PI;
      v = v + 0.5 * extra.y / M_PI;
    }
    int tj = L.tj;
    asm volatile("" : "+v"(tj));
    a.uv[a.perm[tj]] = make_double2(u, v);
  }
  wave_sync();
}

// One wave per target leaf and one wave per workgroup: a finished wave's
// slot is refilled at once (with several waves per workgroup the slot waits
// for the workgroup's slowest leaf).  At least four waves per SIMD: the walk
// is latency-bound and needs the others to hide it.
//
// Workgroups are dealt to the 8 XCDs in turn (workgroup b runs on XCD b mod
// 8); XCD x is given blocks x, x + 8, x + 16, ... of kXcdBlock consecutive
// leaves (neighbours in space), so the tree nodes and sources of one block's
// walks meet in one L2 while the blocks still spread the clustered work
// evenly over the XCDs.  Grid: a multiple of 8 kXcdBlock workgroups.  The
// positions are those of the launch order (lorder: boxes large against the
// near radius first, so the longest walks do not start last).
constexpr int kXcds = 8;
constexpr int kXcdBlock = 64;
__device__ inline int xcd_leaf(int b, int nleaves) {
  const int x = b % kXcds, k = b / kXcds;
  return ((k / kXcdBlock) * kXcds + x) * kXcdBlock + k % kXcdBlock;
}

template <bool MIXED, bool UNI>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void fmm_leaf_kernel(
    FmmArgs a) {
  __shared__ double2 s_loc[1][kFmmMaxP];
  __shared__ double2 s_red[1][64];
  __shared__ typename StageOf<MIXED>::T s_stage[1][stage_slots<MIXED>()];
  __shared__ int64_t s_ck[1][kMaxCk];
  __shared__ int s_stack[1][kStack];
  __shared__ int s_far[1][kFarList];
  __shared__ int32_t s_cku[1][kMaxCk];
  __shared__ ParentBox s_pbox[1];
  const int q = xcd_leaf(blockIdx.x, a.npos);
  if (q >= a.npos) return;
  const int pos = a.pos0 + q;
  const int li = a.lorder ? (int)a.lorder[pos] : pos;
  fmm_process_leaf<MIXED, UNI>(a, li, 0, threadIdx.x, s_loc, s_red, s_stage, s_ck, s_cku, s_stack,
                               s_far, s_pbox);
}

// Parent-level far field (L2L; see ParentCtx): one wave per sibling group g
// whose parent has at least two target-leaf children.  The parent's walk
// opens a node iff every one of those children opens it and lists an opened
// node's child iff it is far from the parent (R_P, the largest near radius of
// the children); one far_contract with the largest order of the children puts
// the parent's scaled local expansion into plocal[g].  Grid-stride over the
// groups, their number read from the device (a launch that can be captured
// once and replayed on rebuilt trees).  pcount (optional): per group its far
// cells, the parent expansion's terms and its target-leaf children (zeros for a
// group without a parent expansion).
template <bool MIXED, bool UNI, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) void parent_local_kernel(
    FmmArgs a, int64_t *__restrict__ pcount) {
  __shared__ double2 s_loc[W][kFmmMaxP];
  __shared__ int s_stack[W][kStack];
  __shared__ int s_far[W][kFarList];
  const int w = W == 1 ? 0 : (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ng = (int)*a.ngroups;
  const bool smooth_on = MIXED && UNI && a.smooth_region > 0.0;
  const double inv_theta = 1.0 / a.theta;
  for (int g = blockIdx.x * W + w; g < ng; g += gridDim.x * W) {
    const int par = rfl(ldro(a.gparent + g));
    NodeRec C4[4];
    double rc[4];
    int pq[4];
    const uint32_t m = par >= 0 ? group_targets(a, g, C4, rc, pq) : 0u;
    if (__builtin_popcount(m) < a.l2l) {
      if (pcount && lane == 0) pcount[3 * (size_t)g] = pcount[3 * (size_t)g + 1] = pcount[3 * (size_t)g + 2] = 0;
      if (a.pfront && lane == 0) a.pfcnt[g] = -1;
      continue;
    }
    const NodeRec NP = ldro(a.nodes + par);
    const double RP = NP.region * 0.70710678118654752;
    double rcutP = 0.0;
    int pP = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if ((m >> q) & 1) {
        rcutP = fmax(rcutP, rc[q]);
        pP = max(pP, pq[q]);
      }
    }
    // the children of one parent share a region; the first target leaf's first
    // blob names the tree (forests)
    const int q0 = __builtin_ctz(m);
    const double region = 0.5 * NP.region, RA = region * 0.70710678118654752;
    const int first0 = (q0 == 0) ? C4[0].first : (q0 == 1) ? C4[1].first : (q0 == 2) ? C4[2].first
                                                                                       : C4[3].first;
    const double2 sPpow = cpow_u(make_double2(NP.region, 0.0), (lane & 15) + 1);
    if (lane < kFmmMaxP) s_loc[w][lane] = make_double2(0.0, 0.0);
    int *stack = s_stack[w];
    const int root = rfl(root_of(a, first0));
    int ns = 1, nfar = 0, nfr = 0;
    int64_t total = 0;
    if (lane == 0) stack[0] = ldro(a.nodes + root).first;
    wave_sync();
    while (ns > 0) {
      const int k = min(min(16, ns), max(1, (kStack - kStackMargin - ns) / 3));
      int c = -1;
      if ((lane >> 2) < k) c = stack[ns - k + (lane >> 2)] + (lane & 3);
      wave_sync();
      ns -= k;
      NodeRec C = {};
      if (c >= 0) C = a.nodes[c];
      bool f = false, op = false;
      if (node_count(C.cnt) > 0) {
        if (far_cell(inv_theta, RP, rcutP, NP, C)) {
          f = true;
        } else if (node_internal(C.cnt)) {
          bool all = true;
#pragma unroll
          for (int q = 0; q < 4; q++)
            if ((m >> q) & 1)
              all = all && leaf_opens(a, smooth_on, inv_theta, RA, rc[q], C4[q].cx, C4[q].cy,
                                      region, c, C);
          op = all;
        }
      }
      const uint64_t fm = __builtin_amdgcn_ballot_w64(f);
      const uint64_t om = __builtin_amdgcn_ballot_w64(op);
      if (op) {
        const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(om >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)om, 0));
        if (ns + below < kStack) stack[ns + below] = C.first;
      }
      ns += __builtin_popcountll(om);
      // the children's start list: each group of four (lanes 4j..4j+3) with
      // a boundary child, the covered and opened ones marked to be skipped
      if (a.pfront) {
        const uint64_t bm = __builtin_amdgcn_ballot_w64(node_count(C.cnt) > 0 && !f && !op);
        const int j4 = lane & ~3;
        const bool head = (lane & 3) == 0 && ((bm >> j4) & 0xFull) != 0;
        const uint64_t hm = __builtin_amdgcn_ballot_w64(head);
        const int nh = __builtin_popcountll(hm);
        if (nfr >= 0 && nfr + nh > a.front_cap) nfr = -1;  // too many: the children walk from the root
        if (nfr >= 0) {
          if (head)
            a.pfront[(size_t)g * kFrontCap + nfr +
                     __builtin_amdgcn_mbcnt_hi((uint32_t)(hm >> 32),
                                               __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0))] =
                (uint32_t)c | ((uint32_t)(((fm | om) >> j4) & 0xFull) << kSkipShift);
          nfr += nh;
        }
      }
      wave_sync();
      if (fm) {
        const int nf = __builtin_popcountll(fm);
        if (nfar + nf > a.far_burst) {
          far_contract<MIXED>(a, NP, NP.region, pP, sPpow, lane, s_far[w], nfar, s_loc[w]);
          nfar = 0;
        }
        if (f)
          s_far[w][nfar + __builtin_amdgcn_mbcnt_hi((uint32_t)(fm >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)fm, 0))] = c;
        nfar += nf;
        total += nf;
        wave_sync();
      }
    }
    if (nfar > 0) far_contract<MIXED>(a, NP, NP.region, pP, sPpow, lane, s_far[w], nfar, s_loc[w]);
    if (lane < kFmmMaxP) a.plocal[(size_t)g * kFmmMaxP + lane] = s_loc[w][lane];
    if (a.pfront && lane == 0) a.pfcnt[g] = nfr;
    if (pcount && lane == 0) {
      pcount[3 * (size_t)g] = total;
      pcount[3 * (size_t)g + 1] = pP;
      pcount[3 * (size_t)g + 2] = __builtin_popcount(m);
    }
    wave_sync();
  }
}

__global__ __launch_bounds__(256) void l2l_setup_kernel(const int32_t *__restrict__ leaves,
                                                        int nleaves, int32_t *__restrict__ nodeli,
                                                        uint32_t *__restrict__ ngroups_dev,
                                                        uint32_t ngroups) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0) *ngroups_dev = ngroups;
  if (i < nleaves) nodeli[leaves[i]] = i;
}

int launch_l2l_setup(const int32_t *d_leaves, int nleaves, int32_t *d_nodeli,
                     uint32_t *d_ngroups, int ngroups, hipStream_t s) {
  l2l_setup_kernel<<<(std::max(nleaves, 1) + 255) / 256, 256, 0, s>>>(d_leaves, nleaves, d_nodeli,
                                                                      d_ngroups, (uint32_t)ngroups);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_parent_local(const FmmArgs &args, bool mixed, int64_t *d_pcount, int nblocks,
                        hipStream_t s) {
  if (!args.l2l) return 0;
  if (args.ngroups_h > 0) {
    // one wave per sibling group, one wave per workgroup (a finished wave's
    // slot is refilled at once)
    const int nb = args.ngroups_h;
    if (mixed && args.sw > 0.f)
      parent_local_kernel<true, true, 1><<<nb, 64, 0, s>>>(args, d_pcount);
    else if (mixed)
      parent_local_kernel<true, false, 1><<<nb, 64, 0, s>>>(args, d_pcount);
    else
      parent_local_kernel<false, false, 1><<<nb, 64, 0, s>>>(args, d_pcount);
  } else {
    // the group count only on the device (graph replay): grid-stride
    if (mixed && args.sw > 0.f)
      parent_local_kernel<true, true, 4><<<nblocks, 256, 0, s>>>(args, d_pcount);
    else if (mixed)
      parent_local_kernel<true, false, 4><<<nblocks, 256, 0, s>>>(args, d_pcount);
    else
      parent_local_kernel<false, false, 4><<<nblocks, 256, 0, s>>>(args, d_pcount);
  }
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// Persistent form: a fixed grid whose waves take target leaves from a work
// counter until the device-side leaf count is exhausted.  Every wave exits
// once the counter passes the count, so the grid always drains.  Nothing in
// the launch depends on host-side counts, so it can be captured in a graph
// and replayed on a rebuilt tree.  One wave per workgroup at four waves per
// SIMD, like the grid form (four waves per workgroup at three per SIMD ran
// 1.1-1.4x slower).
template <bool MIXED, bool UNI>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void fmm_leaf_persistent_kernel(FmmArgs a,
                                                                  const uint32_t *__restrict__ nleaves,
                                                                  uint32_t *__restrict__ counter) {
  __shared__ double2 s_loc[1][kFmmMaxP];
  __shared__ double2 s_red[1][64];
  __shared__ typename StageOf<MIXED>::T s_stage[1][stage_slots<MIXED>()];
  __shared__ int64_t s_ck[1][kMaxCk];
  __shared__ int s_stack[1][kStack];
  __shared__ int s_far[1][kFarList];
  __shared__ int32_t s_cku[1][kMaxCk];
  __shared__ ParentBox s_pbox[1];
  const int w = 0, lane = threadIdx.x;
  const uint32_t total = *nleaves;
  for (;;) {
    uint32_t li = 0;
    if (lane == 0) li = atomicAdd(counter, 1u);
    li = __shfl(li, 0, 64);
    if (li >= total) break;
    if (a.lorder) li = a.lorder[li];
    fmm_process_leaf<MIXED, UNI>(a, (int)li, w, lane, s_loc, s_red, s_stage, s_ck, s_cku,
                                 s_stack, s_far, s_pbox);
  }
}

// Continuation pass: waves take chunk records from a work counter; the
// partial sums of record j go to part[64 j + target].  Exits when the counter
// passes the number of records written by the fused pass.
template <bool MIXED, bool UNI>
__global__ __launch_bounds__(256) void fmm_chunk_kernel(FmmArgs a, uint32_t *__restrict__ counter) {
  __shared__ double2 s_red[4][64];
  __shared__ typename StageOf<MIXED>::T s_stage[4][stage_slots<MIXED>()];
  __shared__ int s_stack[4][kStack];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t total = min(*a.ck_count, (uint32_t)a.ck_cap);
  for (;;) {
    uint32_t j = 0;
    if (lane == 0) j = atomicAdd(counter, 1u);
    j = (uint32_t)rfl((int)__shfl((int)j, 0, 64));
    if (j >= total) break;
    const SplitCk r = ldro(a.ck + j);
    const int li = rfl(r.li);
    if (li < 0) continue;
    const LeafCtx L = leaf_ctx<MIXED, UNI>(a, li, lane);
    double ax = 0.0, ay = 0.0;
    run_chunk<MIXED, UNI>(a, L, li, (int64_t)rfl64((uint64_t)r.cum),
                          (int64_t)rfl64((uint64_t)r.end), lane, s_stage[w], s_stack[w], ax, ay);
    const double2 P = reduce_replicas(L, lane, s_red[w], ax, ay);
    if (L.tact && L.rep == 0) a.part[((size_t)a.ck_base + j) * 64 + L.ti] = P;
  }
}

// Adds the chunk sums of every split leaf (split_list) to its targets, in
// chunk order.
__global__ __launch_bounds__(256) void fmm_split_sum_kernel(FmmArgs a) {
  const int lane = threadIdx.x & 63;
  const uint32_t total = *a.split_count;
  const uint32_t nw = gridDim.x * 4;
  for (uint32_t j = blockIdx.x * 4 + (threadIdx.x >> 6); j < total; j += nw) {
    const int li = a.split_list[j];
    const int2 k = a.leaf_ck[li];
    if (k.y == 0) continue;
    const NodeRec NA = a.nodes[a.leaves[li]];
    if (lane < node_count(NA.cnt)) {
      double sx = 0.0, sy = 0.0;
      // eight loads in flight, then added in chunk order
      const double2 *src = a.part + (size_t)k.x * 64 + lane;
      int i = 0;
      for (; i + 8 <= k.y; i += 8) {
        double2 P[8];
#pragma unroll
        for (int j = 0; j < 8; j++) P[j] = src[(size_t)(i + j) * 64];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          sx += P[j].x;
          sy += P[j].y;
        }
      }
      for (; i < k.y; i++) {
        const double2 P = src[(size_t)i * 64];
        sx += P.x;
        sy += P.y;
      }
      const uint32_t t = a.perm[NA.first + lane];
      double2 uv = a.uv[t];
      uv.x = uv.x + 0.5 * sx / M_PI;
      uv.y = uv.y + 0.5 * sy / M_PI;
      a.uv[t] = uv;
    }
  }
}

__global__ void zero_counters_kernel(uint32_t *c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

// The continuation pass over the records of args' region.
static int launch_chunk_pass(const FmmArgs &args, bool mixed, uint32_t *d_chunk_counter,
                             int nblocks, hipStream_t s) {
  if (mixed && args.sw > 0.f)
    fmm_chunk_kernel<true, true><<<nblocks, 256, 0, s>>>(args, d_chunk_counter);
  else if (mixed)
    fmm_chunk_kernel<true, false><<<nblocks, 256, 0, s>>>(args, d_chunk_counter);
  else
    fmm_chunk_kernel<false, false><<<nblocks, 256, 0, s>>>(args, d_chunk_counter);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

static int launch_split_passes(const FmmArgs &args, bool mixed, uint32_t *d_chunk_counter,
                               int nblocks, hipStream_t s) {
  if (launch_chunk_pass(args, mixed, d_chunk_counter, nblocks, s)) return -1;
  // four split leaves per workgroup, grid-stride over the device-side count
  fmm_split_sum_kernel<<<std::max(1, nblocks / 4), 256, 0, s>>>(args);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_fmm_eval_persistent(const FmmArgs &args, bool mixed, const uint32_t *d_nleaves,
                               uint32_t *d_counter, int nblocks, uint32_t *d_split_ctr,
                               int split_blocks, hipStream_t s) {
  FmmArgs pa = args;
  pa.ngroups_h = 0;  // the launch may be captured and replayed on another tree
  if (launch_parent_local(pa, mixed, nullptr, split_blocks, s)) return -1;
  zero_counters_kernel<<<1, 64, 0, s>>>(d_counter, 1);
  FMM_HIP_TRY(hipGetLastError());
  const bool split = args.split_budget > 0;
  if (split) {
    zero_counters_kernel<<<1, 64, 0, s>>>(d_split_ctr, 5);
    FMM_HIP_TRY(hipGetLastError());
  }
  if (mixed && args.sw > 0.f)
    fmm_leaf_persistent_kernel<true, true><<<4 * nblocks, 64, 0, s>>>(args, d_nleaves, d_counter);
  else if (mixed)
    fmm_leaf_persistent_kernel<true, false><<<4 * nblocks, 64, 0, s>>>(args, d_nleaves, d_counter);
  else
    fmm_leaf_persistent_kernel<false, false><<<4 * nblocks, 64, 0, s>>>(args, d_nleaves, d_counter);
  FMM_HIP_TRY(hipGetLastError());
  if (split) return launch_split_passes(args, mixed, d_split_ctr + 1, split_blocks, s);
  return 0;
}

// Work of the fused kernel, one thread per target leaf with the same walk:
// out[kCountWords*leaf + 0] = far cells (each one M2L into the leaf), + 1 = near
// source-target pairs summed directly, + 2 = (target, smooth cell)
// evaluations, + 3 = the leaf's expansion terms, + 4 = its targets.
__global__ __launch_bounds__(256) void fmm_count_kernel(FmmArgs a, int64_t *__restrict__ out) {
  const int li = blockIdx.x * 256 + threadIdx.x;
  if (li >= a.nleaves) return;
  const int iA = a.leaves[li];
  const NodeRec NA = a.nodes[iA];
  const int cA = node_count(NA.cnt);
  const double RA = NA.region * 0.70710678118654752;
  const double rcut = a.leaf_par ? (double)a.leaf_par[li].y * a.smax : a.rcut;
  const double inv_theta = 1.0 / a.theta;
  const bool smooth_on = a.smooth_region > 0.0;
  // the parent's part (L2L, see ParentCtx): the same rules as the fused walk
  bool pon = false;
  NodeRec NP = {};
  double RP = 0.0, rcutP = 0.0, sx[3] = {0, 0, 0}, sy[3] = {0, 0, 0}, srcut[3] = {0, 0, 0};
  int nsib = 0;
  if (a.l2l && a.gparent[iA >> 2] >= 0) {
    const int g = iA >> 2;
    int nt = 0;
    for (int q = 0; q < 4; q++) {
      const NodeRec C = a.nodes[4 * g + q];
      if (node_internal(C.cnt) || node_count(C.cnt) == 0) continue;
      double rc;
      int pq;
      leaf_rcut_p(a, a.nodeli[4 * g + q], rc, pq);
      rcutP = fmax(rcutP, rc);
      nt++;
      if (q != (iA & 3)) {
        sx[nsib] = C.cx;
        sy[nsib] = C.cy;
        srcut[nsib] = rc;
        nsib++;
      }
    }
    if (nt >= a.l2l) {
      pon = true;
      NP = a.nodes[a.gparent[g]];
      RP = NP.region * 0.70710678118654752;
    }
  }
  int64_t far = 0, pairs = 0, smooth = 0;
  Slots2 sl;
  uint32_t covd = 0;  // bit d: the node expanded at depth d was opened by the parent
  int d = 0, cur = 0, node = root_of(a, NA.first);
  bool expand = true, fl = pon;
  for (;;) {
    if (expand) {
      const int base = a.nodes[node].first;
      uint32_t m = 0;
      for (int q = 0; q < 4; q++) {
        const NodeRec C = a.nodes[base + q];
        if (node_count(C.cnt) == 0) continue;
        if (fl && far_cell(inv_theta, RP, rcutP, NP, C)) continue;  // the parent's
        if (far_cell(inv_theta, RA, rcut, NA, C)) far++;
        else m |= 1u << q;
      }
      d++;
      sl.set(d, m);
      covd = fl ? (covd | (1u << d)) : (covd & ~(1u << d));
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
    const NodeRec Cn = a.nodes[node];
    if (smooth_cell(a, Cn)) {
      smooth += cA;
      expand = false;
    } else if (node_internal(Cn.cnt)) {
      expand = true;
      // the children of Cn are the parent's iff the parent opened Cn: every
      // target-leaf child opens it (A's walk opens it unless it is a whole near
      // subtree taken as one item)
      fl = false;
      if ((covd >> d) & 1u) {
        fl = !(whole_near(inv_theta, RA, rcut, NA, Cn) &&
               (!smooth_on || a.srange[node].x != -2));
        for (int s2 = 0; s2 < nsib && fl; s2++)
          fl = leaf_opens(a, smooth_on, inv_theta, RA, srcut[s2], sx[s2], sy[s2], NA.region, node,
                          Cn);
      }
    } else {
      pairs += (int64_t)cA * node_count(Cn.cnt);
      expand = false;
    }
  }
  int64_t *o = out + (size_t)kCountWords * li;
  o[0] = far;
  o[1] = pairs;
  o[2] = smooth;
  o[3] = a.leaf_par ? min(a.p, max(2, (int)a.leaf_par[li].x)) : a.p;
  o[4] = cA;
}

// ---------------------------------------------------------------------------
// Accurate modes at arbitrary points (not blobs of the tree: e.g. the sheet
// centres of e_velo_mpi.f90:59-70): one lane per point, depth first through
// the tree.  The far test is the target leaf's with the leaf shrunk to the
// point (R_A = 0): a cell is far iff R_C <= theta d and d - R_C >= r_cut, so
// the neglected core term and the truncation obey the same bounds as for the
// blobs.  A far cell's multipole is evaluated at the point,
//   phi += sum_{k<p} a_k w^(k+1),  w = 1 / (z - c_C)   (Horner),
// and near leaves are summed with the regularised kernel, both in FP64 in
// either accurate mode (few points: precision is cheaper than a second path).
// Fixed p and r_cut (learned per-leaf values belong to target leaves).
// ---------------------------------------------------------------------------
__device__ inline void m2p_horner(const double2 *__restrict__ a, int p, double dx, double dy,
                                  double &phr, double &phi) {
  const double r2 = fma(dx, dx, dy * dy);
  double ri = __builtin_amdgcn_rcp(r2);
  ri = ri * fma(-r2, ri, 2.0);
  ri = ri * fma(-r2, ri, 2.0);
  const double wr = dx * ri, wi = -dy * ri;
  double sr = 0.0, si = 0.0;
  for (int k = p - 1; k >= 0; k--) {
    const double2 ak = a[k];
    const double nr = fma(sr, wr, fma(-si, wi, ak.x));
    const double ni = fma(sr, wi, fma(si, wr, ak.y));
    sr = nr;
    si = ni;
  }
  phr += sr * wr - si * wi;
  phi += sr * wi + si * wr;
}

__global__ __launch_bounds__(256) void point_eval_kernel(FmmArgs a, int nt,
                                                         const double2 *__restrict__ txy,
                                                         double2 *__restrict__ uv) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nt) return;
  const double2 z = txy[t];
  NodeRec P = {};
  P.cx = z.x;
  P.cy = z.y;
  const double inv_theta = 1.0 / a.theta;
  double phr = 0.0, phi = 0.0, ax = 0.0, ay = 0.0;
  Slots2 sl;
  int d = 0, cur = 0, node = 0;  // a single tree: root 0, always split
  bool expand = true;
  for (;;) {
    if (expand) {
      const int base = a.nodes[node].first;
      uint32_t m = 0;
      for (int q = 0; q < 4; q++) {
        const NodeRec C = a.nodes[base + q];
        if (node_count(C.cnt) == 0) continue;
        if (far_cell(inv_theta, 0.0, a.rcut, P, C))
          m2p_horner(a.coef + (size_t)(base + q) * a.p, a.p, z.x - C.cx, z.y - C.cy, phr, phi);
        else
          m |= 1u << q;
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
      if (d == 1) {
        done = true;
        break;
      }
      cur = a.gparent[cur >> 2];
      d--;
    }
    if (done) break;
    const int q = 31 - __clz(m);
    sl.set(d, m & ~(1u << q));
    node = (cur & ~3) + q;
    const NodeRec N = a.nodes[node];
    if (node_internal(N.cnt)) {
      expand = true;
    } else {
      const int f = N.first, c = node_count(N.cnt);
      for (int j = f; j < f + c; j++) pair64(z.x, z.y, a.loc64[j], ax, ay);
      expand = false;
    }
  }
  uv[t] = make_double2(0.5 * phi / M_PI + 0.5 * ax / M_PI, 0.5 * phr / M_PI + 0.5 * ay / M_PI);
}

int launch_point_eval(const FmmArgs &args, int nt, const double *d_txy, double *d_uv,
                      hipStream_t s) {
  if (nt <= 0) return 0;
  point_eval_kernel<<<(nt + 255) / 256, 256, 0, s>>>(args, nt,
                                                     reinterpret_cast<const double2 *>(d_txy),
                                                     reinterpret_cast<double2 *>(d_uv));
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_fmm_count(const FmmArgs &args, int64_t *d_out, hipStream_t s) {
  if (args.nleaves <= 0) return 0;
  fmm_count_kernel<<<(args.nleaves + 255) / 256, 256, 0, s>>>(args, d_out);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// The fused pass over launch positions [args.pos0, args.pos0 + args.npos).
static int launch_leaf_grid(const FmmArgs &args, bool mixed, hipStream_t s) {
  if (args.npos <= 0) return 0;
  const int grid = kXcds * kXcdBlock * ((args.npos + kXcds * kXcdBlock - 1) / (kXcds * kXcdBlock));
  if (mixed && args.sw > 0.f)
    fmm_leaf_kernel<true, true><<<grid, 64, 0, s>>>(args);
  else if (mixed)
    fmm_leaf_kernel<true, false><<<grid, 64, 0, s>>>(args);
  else
    fmm_leaf_kernel<false, false><<<grid, 64, 0, s>>>(args);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_fmm_eval(const FmmArgs &args, bool mixed, const uint32_t *d_nleaves,
                    uint32_t *d_split_ctr, int split_blocks, hipStream_t s, hipEvent_t mid,
                    hipStream_t side, hipEvent_t fork, hipEvent_t join, int nfront,
                    hipEvent_t ready, int side_blocks, hipEvent_t ready1) {
  if (args.nleaves <= 0) {
    if (ready) FMM_HIP_TRY(hipStreamWaitEvent(s, ready, 0));
    return 0;
  }
  if (ready1) FMM_HIP_TRY(hipStreamWaitEvent(s, ready1, 0));
  if (launch_parent_local(args, mixed, nullptr, split_blocks, s)) return -1;
  if (ready) FMM_HIP_TRY(hipStreamWaitEvent(s, ready, 0));
  const bool split = args.split_budget > 0;
  if (split) {
    zero_counters_kernel<<<1, 64, 0, s>>>(d_split_ctr, 5);
    FMM_HIP_TRY(hipGetLastError());
  }
  FmmArgs a = args;
  a.pos0 = 0;
  a.npos = args.nleaves;
  a.ck_base = 0;
  if (split && side && fork && join && nfront > 0 && nfront < args.nleaves) {
    // the front leaves (the longest walks; they hold nearly all the split
    // leaves) and their continuation pass on s, concurrently with the other
    // leaves on the side stream: the continuation pass overlaps the fused
    // pass instead of following it on an idle device
    FmmArgs b = a;
    a.npos = nfront;
    b.pos0 = nfront;
    b.npos = args.nleaves - nfront;
    b.ck = args.ck + args.ck_cap;
    b.ck_count = d_split_ctr + 2;
    b.ck_base = args.ck_cap;
    FMM_HIP_TRY(hipEventRecord(fork, s));
    FMM_HIP_TRY(hipStreamWaitEvent(side, fork, 0));
    if (launch_leaf_grid(a, mixed, s) || launch_leaf_grid(b, mixed, side)) return -1;
    if (launch_chunk_pass(a, mixed, d_split_ctr + 1, split_blocks, s)) return -1;
    FMM_HIP_TRY(hipEventRecord(join, side));
    FMM_HIP_TRY(hipStreamWaitEvent(s, join, 0));
    if (mid) FMM_HIP_TRY(hipEventRecord(mid, s));
    return launch_split_passes(b, mixed, d_split_ctr + 3, std::max(1, side_blocks), s);
  }
  if (launch_leaf_grid(a, mixed, s)) return -1;
  if (mid) FMM_HIP_TRY(hipEventRecord(mid, s));
  if (split) return launch_split_passes(a, mixed, d_split_ctr + 1, split_blocks, s);
  return 0;
}

}  // namespace fmm2d
