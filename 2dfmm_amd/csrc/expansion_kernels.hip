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
template <int NV>
__global__ __launch_bounds__(256) void s2m_kernel(const NodeRec *__restrict__ nodes, int nnodes,
                                                  const double *__restrict__ sx,
                                                  const double *__restrict__ sy,
                                                  const double *__restrict__ sg, int nterms,
                                                  double2 *__restrict__ coef) {
  const int node = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (node >= nnodes) return;
  const NodeRec r = nodes[node];
  const int c = node_count(r.cnt);
  if (node_internal(r.cnt) || c == 0) return;
  // v[2k], v[2k+1] = cir (z - z_c)^k (real, imaginary) of this lane's blob;
  // after the reduce-scatter lane l holds the wave's sum of v[l * NV / 64]
  double acc = 0.0;
  for (int base = 0; base < c; base += 64) {
    const int e = base + lane;
    const bool ok = e < c;
    const double tr = ok ? sx[r.first + e] - r.cx : 0.0;
    const double ti = ok ? sy[r.first + e] - r.cy : 0.0;
    const double g = ok ? sg[r.first + e] : 0.0;
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
}

int launch_s2m(const NodeRec *d_nodes, int nnodes, const double *sx, const double *sy,
               const double *sg, int nterms, double2 *d_coef, hipStream_t s) {
  if (nterms <= 16)
    s2m_kernel<32><<<(nnodes + 3) / 4, 256, 0, s>>>(d_nodes, nnodes, sx, sy, sg, nterms, d_coef);
  else
    s2m_kernel<64><<<(nnodes + 3) / 4, 256, 0, s>>>(d_nodes, nnodes, sx, sy, sg, nterms, d_coef);
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

int launch_m2m_level(const NodeRec *d_nodes, const int32_t *d_parents, int nparents, int nterms,
                     double2 *d_coef, hipStream_t s) {
  if (nparents <= 0) return 0;
  m2m_kernel<<<nparents, 64, 0, s>>>(d_nodes, d_parents, nparents, nterms, d_coef);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace fmm2d
