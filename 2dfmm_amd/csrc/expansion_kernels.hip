// expansion_kernels.hip -- multipole expansions on the GPU (gfx950).
//
//   S2M  a_k = sum_j cir_j (z_j - z_c)^k, k < nterms, for every non-empty
//        leaf (fmm.c:18-33).  One wave per leaf, one blob per lane; the powers
//        are carried by repeated multiplication and the 2 nterms sums are one
//        reduce-scatter across the wave (lane 4k: Re a_k, lane 4k+2: Im a_k).
//   M2M  b_n = sum over non-empty children, quadrant order 3,2,1,0, of
//        sum_{k<=n} a_k C(n,k) (z_child - z_parent)^(n-k)   (fmm.c:35-58).
//        One wave per parent, lane n builds b_n; the powers of the shift are
//        shared through LDS.  Launched level by level, deepest first; the
//        level-1 nodes are not shifted into the root (fmm.c:41).
#include <hip/hip_runtime.h>

#include "fmm2d_internal.h"

namespace fmm2d {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One halving step of a reduce-scatter across the wave: lanes whose bit `o`
// is clear keep the lower half of v, the others the upper half, each adding
// its partner's copy of the kept half.
template <int H, int NV>
__device__ inline void halve(double (&v)[NV], int lane, int o) {
  const bool up = (lane & o) != 0;
#pragma unroll
  for (int i = 0; i < H; i++) {
    const double keep = up ? v[H + i] : v[i];
    const double give = up ? v[i] : v[H + i];
    v[i] = keep + __shfl_xor(give, o, 64);
  }
}

// NV = 2 x (terms rounded up to 16 or 32) values per lane.
// Nodes that get no expansion otherwise -- empty leaves and the roots (never
// shifted into, fmm.c:41) -- get zeros here, so no clearing pass is needed.
// leaf_abs (optional): sum of |cir| over each leaf's blobs (zero for empty
// leaves and roots), summed per lane and then across the wave in a fixed
// order.
template <int NV>
__global__ __launch_bounds__(256) void s2m_kernel(const NodeRec *__restrict__ nodes, int nnodes,
                                                  int nroot_slots, const double *__restrict__ sx,
                                                  const double *__restrict__ sy,
                                                  const double *__restrict__ sg, int nterms,
                                                  double2 *__restrict__ coef,
                                                  double *__restrict__ leaf_abs) {
  const int node = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (node >= nnodes) return;
  const NodeRec r = nodes[node];
  const int c = node_count(r.cnt);
  if (node < nroot_slots || (!node_internal(r.cnt) && c == 0)) {
    if (lane < nterms) coef[(size_t)node * nterms + lane] = make_double2(0.0, 0.0);
    if (leaf_abs && lane == 0) leaf_abs[node] = 0.0;
    return;
  }
  if (node_internal(r.cnt)) return;
  double gabs = 0.0;
  // v[2k], v[2k+1] = cir (z - z_c)^k (real, imaginary) of this lane's blob;
  // after the reduce-scatter lane l holds the wave's sum of v[l * NV / 64]
  double acc = 0.0;
  for (int base = 0; base < c; base += 64) {
    const int e = base + lane;
    const bool ok = e < c;
    const double tr = ok ? sx[r.first + e] - r.cx : 0.0;
    const double ti = ok ? sy[r.first + e] - r.cy : 0.0;
    const double g = ok ? sg[r.first + e] : 0.0;
    gabs += fabs(g);
    double v[NV];
    double pr = 1.0, pi = 0.0;
#pragma unroll
    for (int k = 0; k < NV / 2; k++) {
      v[2 * k] = k < nterms ? g * pr : 0.0;
      v[2 * k + 1] = k < nterms ? g * pi : 0.0;
      const double nr = pr * tr - pi * ti;
      const double ni = pr * ti + pi * tr;
      pr = nr;
      pi = ni;
    }
    halve<NV / 2>(v, lane, 32);
    halve<NV / 4>(v, lane, 16);
    halve<NV / 8>(v, lane, 8);
    halve<NV / 16>(v, lane, 4);
    halve<NV / 32>(v, lane, 2);
    if (NV == 64) {
      halve<1>(v, lane, 1);
      acc += v[0];
    } else {
      acc += v[0] + __shfl_xor(v[0], 1, 64);
    }
  }
  const int j = NV == 64 ? lane : lane >> 1, k = j >> 1;
  if ((NV == 64 || (lane & 1) == 0) && k < nterms)
    reinterpret_cast<double *>(coef)[2 * ((size_t)node * nterms + k) + (j & 1)] = acc;
  if (leaf_abs) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) gabs += __shfl_xor(gabs, o, 64);
    if (lane == 0) leaf_abs[node] = gabs;
  }
}

// nterms <= 16: one node per half-wave (a leaf holds 29 blobs on average at
// leaf size 64, so a whole wave per leaf idled half of it).  The 32 lanes
// go through the leaf's blobs 32 at a time; the reduce-scatter halves stay
// inside the 32 lanes and leave lane j with the sum of value j (terms k =
// j / 2, real or imaginary part j % 2).
__global__ __launch_bounds__(256) void s2m_half_kernel(const NodeRec *__restrict__ nodes, int nnodes,
                                                       int nroot_slots, const double *__restrict__ sx,
                                                       const double *__restrict__ sy,
                                                       const double *__restrict__ sg, int nterms,
                                                       double2 *__restrict__ coef,
                                                       double *__restrict__ leaf_abs) {
  constexpr int NV = 32;
  const int node = blockIdx.x * 8 + (threadIdx.x >> 5);
  const int lane = threadIdx.x & 31;
  if (node >= nnodes) return;
  const NodeRec r = nodes[node];
  const int c = node_count(r.cnt);
  if (node < nroot_slots || (!node_internal(r.cnt) && c == 0)) {
    if (lane < nterms) coef[(size_t)node * nterms + lane] = make_double2(0.0, 0.0);
    if (leaf_abs && lane == 0) leaf_abs[node] = 0.0;
    return;
  }
  if (node_internal(r.cnt)) return;
  double gabs = 0.0, acc = 0.0;
  for (int base = 0; base < c; base += 32) {
    const int e = base + lane;
    const bool ok = e < c;
    const double tr = ok ? sx[r.first + e] - r.cx : 0.0;
    const double ti = ok ? sy[r.first + e] - r.cy : 0.0;
    const double g = ok ? sg[r.first + e] : 0.0;
    gabs += fabs(g);
    double v[NV];
    double pr = 1.0, pi = 0.0;
#pragma unroll
    for (int k = 0; k < NV / 2; k++) {
      v[2 * k] = k < nterms ? g * pr : 0.0;
      v[2 * k + 1] = k < nterms ? g * pi : 0.0;
      const double nr = fma(pr, tr, -(pi * ti));
      const double ni = fma(pr, ti, pi * tr);
      pr = nr;
      pi = ni;
    }
    halve<NV / 2>(v, lane, 16);
    halve<NV / 4>(v, lane, 8);
    halve<NV / 8>(v, lane, 4);
    halve<NV / 16>(v, lane, 2);
    halve<NV / 32>(v, lane, 1);
    acc += v[0];
  }
  const int k = lane >> 1;
  if (k < nterms) reinterpret_cast<double *>(coef)[2 * ((size_t)node * nterms + k) + (lane & 1)] = acc;
  if (leaf_abs) {
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) gabs += __shfl_xor(gabs, o, 64);
    if (lane == 0) leaf_abs[node] = gabs;
  }
}

int launch_s2m(const NodeRec *d_nodes, int nnodes, int nroot_slots, const double *sx,
               const double *sy, const double *sg, int nterms, double2 *d_coef, double *d_leaf_abs,
               hipStream_t s) {
  if (nterms <= 16)
    s2m_half_kernel<<<(nnodes + 7) / 8, 256, 0, s>>>(d_nodes, nnodes, nroot_slots, sx, sy, sg,
                                                     nterms, d_coef, d_leaf_abs);
  else
    s2m_kernel<64><<<(nnodes + 3) / 4, 256, 0, s>>>(d_nodes, nnodes, nroot_slots, sx, sy, sg,
                                                    nterms, d_coef, d_leaf_abs);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// d^j by binary exponentiation
__device__ inline double2 cpow_int(double dr, double di, int j) {
  double rr = 1.0, ri = 0.0;
  double br = dr, bi = di;
  while (j > 0) {
    if (j & 1) {
      const double t = rr * br - ri * bi;
      ri = rr * bi + ri * br;
      rr = t;
    }
    j >>= 1;
    if (j) {
      const double t = br * br - bi * bi;
      bi = br * bi + bi * br;
      br = t;
    }
  }
  return make_double2(rr, ri);
}

// One wave (block of 64) per parent node.
__global__ __launch_bounds__(64) void m2m_kernel(const NodeRec *__restrict__ nodes,
                                                 const int32_t *__restrict__ parents, int nparents,
                                                 int nterms, double2 *__restrict__ coef) {
  __shared__ double2 dpow[kMaxTerms];
  __shared__ double2 ach[kMaxTerms];
  const int pi_ = blockIdx.x;
  if (pi_ >= nparents) return;
  const int lane = threadIdx.x;
  const int par = parents[pi_];
  const NodeRec P = nodes[par];
  double br = 0.0, bi = 0.0;
  for (int q = 3; q >= 0; q--) {
    const int ch = P.first + q;
    const NodeRec C = nodes[ch];
    if (node_count(C.cnt) < 1) continue;
    // cpow(-(p_pos - c_pos), n - k) of fmm.c:50
    const double dr = -(P.cx - C.cx), di = -(P.cy - C.cy);
    if (lane < nterms) {
      dpow[lane] = cpow_int(dr, di, lane);
      ach[lane] = coef[(size_t)ch * nterms + lane];
    }
    __syncthreads();
    if (lane < nterms) {
      const int n = lane;
      double cr = 0.0, ci = 0.0;
      double binom = 1.0;
      for (int k = 0; k <= n; k++) {
        if (k > 0) binom = binom * (double)(n - k + 1) / (double)k;
        const double2 a = ach[k];
        const double xr = a.x * binom, xi = a.y * binom;
        const double2 d = dpow[n - k];
        cr += xr * d.x - xi * d.y;
        ci += xr * d.y + xi * d.x;
      }
      br += cr;
      bi += ci;
    }
    __syncthreads();
  }
  if (lane < nterms) coef[(size_t)par * nterms + lane] = make_double2(br, bi);
}

// Up to 16 terms: the four children at once, lane = 16 child + n; their
// terms are then added in the order q = 3 .. 0 of m2m_kernel.
__global__ __launch_bounds__(64) void m2m4_kernel(const NodeRec *__restrict__ nodes,
                                                  const int32_t *__restrict__ parents,
                                                  int nparents, int nterms,
                                                  double2 *__restrict__ coef) {
  __shared__ double2 dpow[4][16];
  __shared__ double2 ach[4][16];
  const int pi_ = blockIdx.x;
  if (pi_ >= nparents) return;
  const int lane = threadIdx.x, q = lane >> 4, n = lane & 15;
  const int par = parents[pi_];
  const NodeRec P = nodes[par];
  const int ch = P.first + q;
  const NodeRec C = nodes[ch];
  const bool has = node_count(C.cnt) >= 1;
  if (n < nterms) {
    // cpow(-(p_pos - c_pos), n - k) of fmm.c:50
    dpow[q][n] = cpow_int(-(P.cx - C.cx), -(P.cy - C.cy), n);
    ach[q][n] = has ? coef[(size_t)ch * nterms + n] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  double cr = 0.0, ci = 0.0;
  if (has && n < nterms) {
    double binom = 1.0;
    for (int k = 0; k <= n; k++) {
      if (k > 0) binom = binom * (double)(n - k + 1) / (double)k;
      const double2 a = ach[q][k];
      const double xr = a.x * binom, xi = a.y * binom;
      const double2 d = dpow[q][n - k];
      cr += xr * d.x - xi * d.y;
      ci += xr * d.y + xi * d.x;
    }
  }
  // children 3, 2, 1, 0 in turn
  double br = 0.0, bi = 0.0;
  for (int qq = 3; qq >= 0; qq--) {
    br += __shfl(cr, 16 * qq + n, 64);
    bi += __shfl(ci, 16 * qq + n, 64);
  }
  if (q == 0 && n < nterms) coef[(size_t)par * nterms + n] = make_double2(br, bi);
}

int launch_m2m_level(const NodeRec *d_nodes, const int32_t *d_parents, int nparents, int nterms,
                     double2 *d_coef, hipStream_t s) {
  if (nparents <= 0) return 0;
  if (nterms <= 16) {
    m2m4_kernel<<<nparents, 64, 0, s>>>(d_nodes, d_parents, nparents, nterms, d_coef);
    FMM_HIP_TRY(hipGetLastError());
    return 0;
  }
  m2m_kernel<<<nparents, 64, 0, s>>>(d_nodes, d_parents, nparents, nterms, d_coef);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// ---- smooth cells ----------------------------------------------------------
// A cell small against the core radius sigma is summed for every target by a
// Taylor expansion of the regularised kernel about the cell centre c.  In core
// radii (z = (x - x_j) / sigma, complex) blob j adds cir_j f(z) / sigma to
// F, the velocity being (u, v) = (Im F, Re F) / (2 pi), with f(z) = (1 - e^{-z zb}) / z,
// zb = conj z, an entire function of (z, zb).  With Z = (x - c) / sigma and
// d_j = (c - x_j) / sigma,
//   sum_j cir_j f(Z + d_j) = sum_{m+n<=q} T_mn(Z) mu_mn,
//   T_mn = d^m/dz^m d^n/dzb^n f,   mu_mn = sum_j cir_j d_j^m conj(d_j)^n / (m! n!),
// the error falling like (|d| / sigma)^(q+1).  Closed forms (E = e^{-z zb}):
//   T_m0 = zb^(m+1) phi^(m)(z zb),  phi(w) = (1 - e^-w) / w    (f = zb phi),
//   T_mn = E H_{m,n-1} for n >= 1,   since d/dzb f = E, and
//   H_mk E = d^m/dz^m d^k/dzb^k E,  H_mk = sum_j C(m,j) k!/(k-j)! (-1)^(m+k-j) z^(k-j) zb^(m-j),
// so the n >= 1 part is E P(z, zb) with P = sum_{i+j<=q-1} c_ij z^i zb^j,
//   c_ij = sum_l mu_{j+l, i+l+1} C(j+l, l) (i+l)!/i! (-1)^(i+j+l).
// The coefficients are stored for the scaled coordinates of the evaluation,
// Z = lam Z' with lam = sqrt(ln 2) (Z' in units of sigma / sqrt(log2 e), where
// e^{-|Z|^2} = 2^{-|Z'|^2}), and divided by sigma: nu_m lam^(m+1) / sigma for
// m = q .. 0, then P in real monomials, P = sum_{a+b<=q-1} d_ab x^a y^b
// (z = x + i y), so that Horner steps multiply by real x and y (half the
// work of complex steps):
//   d_ab = i^b sum_{i+j=a+b} c_ij sum_{k+l=b} (-1)^l C(i,k) C(j,l)
// from (x + iy)^i (x - iy)^j, stored as d_ab lam^(a+b) / sigma in Horner
// order (b = q-1 .. 0, a = q-1-b .. 0).
__device__ inline int smooth_tri(int m) { return m * (kSmoothQ + 1) - m * (m - 1) / 2; }
// storage slot of term (i, j) of the n >= 1 part
__device__ inline int smooth_slot(int i, int j) {
  constexpr int Q = kSmoothQ;
  return Q + 1 + (Q - 1 - j) * (Q - j) / 2 + (Q - 1 - j - i);
}

// k! for k <= kSmoothQ (exact in FP64: 12! < 2^53; no formula below needs
// more), a constant table rather than a per-lane array (which would live in
// scratch memory)
struct SmoothFact {
  double v[kSmoothQ + 1];
};
constexpr SmoothFact make_smooth_fact() {
  SmoothFact t{};
  t.v[0] = 1.0;
  for (int k = 1; k <= kSmoothQ; k++) t.v[k] = t.v[k - 1] * (double)k;
  return t;
}
__constant__ SmoothFact kSmoothFact = make_smooth_fact();
// C(n, k) and the falling products (i + l)! / i! for n, i + l <= kSmoothQ:
// integers below 2^29, exact in FP64, so the coefficient formulas need no
// division
struct SmoothInt {
  double binom[kSmoothQ + 1][kSmoothQ + 1];
  double rise[kSmoothQ + 1][kSmoothQ + 1];  // rise[i][l] = (i + l)! / i!
};
constexpr SmoothInt make_smooth_int() {
  SmoothInt t{};
  for (int n = 0; n <= kSmoothQ; n++) {
    t.binom[n][0] = 1.0;
    for (int k = 1; k <= n; k++) t.binom[n][k] = t.binom[n][k - 1] * (double)(n - k + 1) / (double)k;
  }
  for (int i = 0; i <= kSmoothQ; i++) {
    t.rise[i][0] = 1.0;
    for (int l = 1; i + l <= kSmoothQ; l++) t.rise[i][l] = t.rise[i][l - 1] * (double)(i + l);
  }
  return t;
}
__constant__ SmoothInt kSmoothInt = make_smooth_int();

// Per-cell constants of the coefficient transform, so that no thread forms
// them in a loop: lam^k (k <= kSmoothQ + 1, lam = sqrt(ln 2), by repeated
// products) and, for every term (a, b) of the n >= 1 part in storage order,
// the integer weights w_i = sum_l (-1)^l C(i, b-l) C(a+b-i, l) (i <= a+b) of
// d_ab = i^b sum_i w_i c_{i, a+b-i}.
struct SmoothConst {
  double lam[kSmoothQ + 2];
  double w[kSmoothC - kSmoothQ - 1][kSmoothQ];
  int8_t ta[kSmoothC - kSmoothQ - 1], tb[kSmoothC - kSmoothQ - 1];
};
constexpr SmoothConst make_smooth_const() {
  constexpr int Q = kSmoothQ;
  SmoothConst t{};
  const SmoothInt bi = make_smooth_int();
  t.lam[0] = 1.0;
  for (int k = 1; k <= Q + 1; k++) t.lam[k] = t.lam[k - 1] * 0.83255461115769775635;
  for (int o = 0; o < kSmoothC - Q - 1; o++) {
    int k = o, j = Q - 1;
    while (k >= Q - j) {
      k -= Q - j;
      j--;
    }
    const int a = Q - 1 - j - k, b = j, n = a + b;
    t.ta[o] = (int8_t)a;
    t.tb[o] = (int8_t)b;
    for (int i = 0; i <= n; i++) {
      const int jj = n - i;
      double w = 0.0;
      for (int l = (b - i > 0 ? b - i : 0); l <= (jj < b ? jj : b); l++) {
        const double v = bi.binom[i][b - l] * bi.binom[jj][l];
        w += (l & 1) ? -v : v;
      }
      t.w[o][i] = w;
    }
  }
  return t;
}
__constant__ SmoothConst kSmoothConst = make_smooth_const();

// Two waves per node (one thread per moment: (q+1)(q+2)/2 = 91 of them).
// The node's blobs go through in chunks of 32: thread j forms cir_j d_j^m and
// conj(d_j)^n (m, n <= q) of blob j in LDS, then thread t < (q+1)(q+2)/2 adds
// the products of term t = (m, n) over the chunk (FP64).  Thread k < kSmoothC
// then forms output coefficient k.
constexpr int kSmoothCoefThreads = 128;
static_assert((kSmoothQ + 1) * (kSmoothQ + 2) / 2 <= kSmoothCoefThreads &&
                  2 * kSmoothF4 <= kSmoothCoefThreads,
              "one thread per moment and per output value");
__global__ __launch_bounds__(kSmoothCoefThreads) void smooth_coef_kernel(const NodeRec *__restrict__ nodes,
                                                         int node0, int nnodes,
                                                         const int32_t *__restrict__ nfirst,
                                                         const double *__restrict__ sx,
                                                         const double *__restrict__ sy,
                                                         const double *__restrict__ sg,
                                                         const double *__restrict__ smax,
                                                         double eta, int smin,
                                                         float4 *__restrict__ out) {
  constexpr int Q = kSmoothQ;
  constexpr int NT = (Q + 1) * (Q + 2) / 2;
  __shared__ double2 mu[NT];
  __shared__ double2 cz[kSmoothC];   // c_ij lam^(i+j) in storage order
  constexpr int B = 32;               // blobs per chunk
  // rows padded by one entry: the lanes of one step read rows m (and n) at
  // the same column, which would otherwise fall into the same LDS banks
  __shared__ double2 pm[Q + 1][B + 1];  // cir d^m
  __shared__ double2 pn[Q + 1][B + 1];  // conj(d)^n
  const int node = node0 + blockIdx.x;
  if (node >= nnodes) return;
  // one core radius only (smax[0] = largest, smax[1] = smallest)
  if (smax[0] != smax[1]) return;
  const double region = 2.0 * eta * smax[0], inv_sigma = 1.0 / smax[0];
  const NodeRec C = nodes[node];
  const int cnt = node_count(C.cnt);
  if (!(C.region <= region && 2.0 * C.region > region && cnt >= smin)) return;
  const int lane = threadIdx.x;
  const double *fact = kSmoothFact.v;
  int m = 0, n = 0;
  if (lane < NT) {
    while (lane >= smooth_tri(m + 1)) m++;
    n = lane - smooth_tri(m);
  }
  const int f = nfirst[node];
  double ar = 0.0, ai = 0.0;
  // this lane's blob of the chunk, loaded a chunk ahead (the loads of the
  // next chunk overlap this chunk's sums)
  double bx = 0.0, by = 0.0, bg = 0.0;
  if (lane < min(B, cnt)) {
    bx = sx[f + lane];
    by = sy[f + lane];
    bg = sg[f + lane];
  }
  for (int j0 = 0; j0 < cnt; j0 += B) {
    const int nb = min(B, cnt - j0);
    if (lane < nb) {
      // d^k by repeated products; cir d^k and conj(d)^k = conj(d^k) from it
      const double dr = (C.cx - bx) * inv_sigma, di = (C.cy - by) * inv_sigma;
      double qr = 1.0, qi = 0.0;
      for (int k = 0; k <= Q; k++) {
        pm[k][lane] = make_double2(bg * qr, bg * qi);
        pn[k][lane] = make_double2(qr, -qi);
        const double t = fma(qr, dr, -(qi * di));  // times d
        qi = fma(qr, di, qi * dr);
        qr = t;
      }
    }
    __syncthreads();
    if (lane < min(B, cnt - j0 - B)) {
      const int j = f + j0 + B + lane;
      bx = sx[j];
      by = sy[j];
      bg = sg[j];
    }
    if (lane < NT) {
      for (int j = 0; j < nb; j++) {
        const double2 u = pm[m][j], v = pn[n][j];
        ar = fma(u.x, v.x, fma(-u.y, v.y, ar));
        ai = fma(u.x, v.y, fma(u.y, v.x, ai));
      }
    }
    __syncthreads();
  }
  if (lane < NT) {
    const double s = 1.0 / (fact[m] * fact[n]);
    mu[lane] = make_double2(ar * s, ai * s);
  }
  __syncthreads();
  double vr = 0.0, vi = 0.0;
  int ti = 0, tj = 0;  // (i, j) of the n >= 1 term of this lane
  {
    if (lane <= Q) {
      const int mm = Q - lane;
      const double2 v = mu[smooth_tri(mm)];
      const double sc = kSmoothConst.lam[mm + 1];
      vr = v.x * sc;
      vi = v.y * sc;
    } else if (lane < kSmoothC) {
      const int i = kSmoothConst.ta[lane - (Q + 1)], j = kSmoothConst.tb[lane - (Q + 1)];
      for (int l = 0; i + j + 2 * l + 1 <= Q; l++) {
        const double2 v = mu[smooth_tri(j + l) + i + l + 1];
        double c = kSmoothInt.binom[j + l][l] * kSmoothInt.rise[i][l];  // C(j+l, l) (i+l)!/i!
        if ((i + j + l) & 1) c = -c;
        vr += v.x * c;
        vi += v.y * c;
      }
      const double sc = kSmoothConst.lam[i + j];
      vr *= sc;
      vi *= sc;
      ti = i;
      tj = j;
    }
  }
  if (lane < kSmoothC) cz[lane] = make_double2(vr, vi);
  __syncthreads();
  if (lane > Q && lane < kSmoothC) {
    // d_ab (a = ti, b = tj): the terms c_ij with i + j = a + b
    const int a = ti, b = tj, n = a + b;
    const double *wt = kSmoothConst.w[lane - (Q + 1)];
    double sr = 0.0, si = 0.0;
    for (int i = 0; i <= n; i++) {
      const int j = n - i;
      const double w = wt[i];  // sum_l (-1)^l C(i, b-l) C(j, l)
      const double2 v = cz[smooth_slot(i, j)];
      sr += w * v.x;
      si += w * v.y;
    }
    switch (b & 3) {  // times i^b
      case 0: vr = sr; vi = si; break;
      case 1: vr = -si; vi = sr; break;
      case 2: vr = -sr; vi = -si; break;
      default: vr = si; vi = -sr; break;
    }
  }
  if (lane <= 2 * (kSmoothF4 - 1) + 1)
    reinterpret_cast<float2 *>(out + (size_t)node * kSmoothF4)[lane] =
        make_float2((float)(vr * inv_sigma), (float)(vi * inv_sigma));
}

int launch_smooth_coef(const NodeRec *d_nodes, int node0, int node1, const int32_t *d_nfirst,
                       const double *sx, const double *sy, const double *sg,
                       const double *d_smax, double eta, int smin, float4 *d_out, hipStream_t s) {
  if (node1 <= node0) return 0;
  smooth_coef_kernel<<<node1 - node0, kSmoothCoefThreads, 0, s>>>(d_nodes, node0, node1, d_nfirst, sx, sy, sg,
                                                  d_smax, eta, smin, d_out);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// flag[n] = 1: a leaf above the smooth level holds blobs of node n's
// subtree; flag[nnodes + n] = 1: a smooth cell (smooth level, >= smin blobs)
// lies in it.  Every such node marks its ancestors (each writer stores 1).
__global__ __launch_bounds__(256) void smooth_mark_kernel(const NodeRec *__restrict__ nodes,
                                                          const int32_t *__restrict__ gparent,
                                                          int nnodes, int nroot_slots, double region,
                                                          int smin, uint8_t *__restrict__ flag) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nnodes) return;
  const NodeRec C = nodes[i];
  const int cnt = node_count(C.cnt);
  uint8_t *f = nullptr;
  if (!node_internal(C.cnt) && cnt > 0 && C.region > region)
    f = flag;
  else if (C.region <= region && 2.0 * C.region > region && cnt >= smin)
    f = flag + nnodes;
  if (!f) return;
  int n = i;
  while (n >= nroot_slots) {
    n = gparent[n >> 2];
    if (n < 0) break;
    f[n] = 1;
  }
}

__global__ __launch_bounds__(256) void smooth_range_kernel(const NodeRec *__restrict__ nodes,
                                                           int nnodes, double region,
                                                           const uint8_t *__restrict__ flag,
                                                           int2 *__restrict__ range) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nnodes) return;
  const NodeRec C = nodes[i];
  int2 r = make_int2(-1, -1);
  if (node_internal(C.cnt) && C.region > region && flag[nnodes + i] && flag[i])
    r = make_int2(-2, -2);
  else if (node_internal(C.cnt) && C.region > region && flag[nnodes + i]) {
    // leftmost and rightmost non-empty paths down to the smooth level (all
    // internal above it: no leaf there holds blobs); a path that meets a
    // leaf leaves the node to be opened
    int lo = i, hi = i;
    NodeRec a = C, b = C;
    bool ok = true;
    while (ok && a.region > region) {
      ok = node_internal(a.cnt) && a.first >= 0 && a.first + 3 < nnodes;
      if (!ok) break;
      int k = 0;
      while (k < 3 && node_count(nodes[a.first + k].cnt) == 0) k++;
      lo = a.first + k;
      a = nodes[lo];
    }
    while (ok && b.region > region) {
      ok = node_internal(b.cnt) && b.first >= 0 && b.first + 3 < nnodes;
      if (!ok) break;
      int k = 3;
      while (k > 0 && node_count(nodes[b.first + k].cnt) == 0) k--;
      hi = b.first + k;
      b = nodes[hi];
    }
    r = ok ? make_int2(lo, hi + 1) : make_int2(-2, -2);
  }
  range[i] = r;
}

int launch_smooth_ranges(const NodeRec *d_nodes, const int32_t *d_gparent, int nnodes,
                         int nroot_slots, double region, int smin, uint8_t *d_flag, int2 *d_range,
                         hipStream_t s) {
  if (nnodes <= 0) return 0;
  const int nb = (nnodes + 255) / 256;
  FMM_HIP_TRY(hipMemsetAsync(d_flag, 0, 2 * (size_t)nnodes, s));
  smooth_mark_kernel<<<nb, 256, 0, s>>>(d_nodes, d_gparent, nnodes, nroot_slots, region, smin,
                                        d_flag);
  FMM_HIP_TRY(hipGetLastError());
  smooth_range_kernel<<<nb, 256, 0, s>>>(d_nodes, nnodes, region, d_flag, d_range);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// Work units of the nodes of the smooth level (the others 0): a smooth cell
// kSmoothUnits, a small leaf its blobs + kStepUnits; then their prefix sums.
__global__ __launch_bounds__(256) void smooth_units_kernel(const NodeRec *__restrict__ nodes,
                                                           int nnodes,
                                                           const double *__restrict__ smax,
                                                           double eta, int smin,
                                                           uint32_t *__restrict__ u) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nnodes) return;
  const double region = 2.0 * eta * smax[0];
  const NodeRec C = nodes[i];
  const int cnt = node_count(C.cnt);
  uint32_t v = 0;
  if (C.region <= region && 2.0 * C.region > region && cnt > 0)
    v = cnt >= smin ? (uint32_t)kSmoothUnits : (uint32_t)(cnt + kStepUnits);
  u[i] = v;
}

int launch_smooth_units(const NodeRec *d_nodes, int nnodes, const double *d_smax, double eta,
                        int smin, uint32_t *d_tmp, uint32_t *d_out, hipStream_t s) {
  if (nnodes <= 0) return 0;
  smooth_units_kernel<<<(nnodes + 255) / 256, 256, 0, s>>>(d_nodes, nnodes, d_smax, eta, smin,
                                                           d_tmp);
  FMM_HIP_TRY(hipGetLastError());
  return scan_u32(d_tmp, d_out, nnodes, d_out + nnodes, s);
}

}  // namespace fmm2d
