// learned_kernels.hip -- learned per-leaf expansion order and near-field radius.
//
// Features of each target leaf A (8, all scale-free):
//   f0 log2(root side / side_A) / 16            depth
//   f1 log2(side_A / sigma_max) / 8             cell size against the core
//   f2 count_A / leaf capacity                  occupancy
//   f3 log10(sum|cir|_A / (count_A mean|cir|)) / 2   circulation strength
//   f4 |a_0| / sum|cir|_A                       sign coherence of the leaf
//   f5 |a_1| / (sum|cir|_A side_A)              dipole moment norm
//   f6 log2(count_A / (side_A / sigma)^2 + 1e-3) / 8  blobs per core area
//   f7 log2(count_parent / count_A) / 4         density contrast
// MLP 8 -> 64 -> 64 -> 2 (ReLU, FP32), weights packed as
//   W1[64][8] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]   (4866 floats)
// Outputs (normalised): p = 11 + 5 y0, cut = 3 + 2 y1 (core radii).
// Training: per-sample forward/backward in one thread, weight gradients by
// one thread per weight over the batch, Adam.  Everything FP32.
#include <hip/hip_runtime.h>

#include "fmm2d_internal.h"

namespace fmm2d {

constexpr int kIn = 8, kH = 64, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kOut * kH, kNW = oB3 + kOut;
static_assert(kNW == kMlpWeights, "weight count");

// One thread per target leaf (the features' logarithms and roots run on all
// lanes; the leaf's |cir| are summed in blob order).
__global__ __launch_bounds__(256) void leaf_features_kernel(const FmmArgs a, int leaf_cap,
                                                            const double *__restrict__ leaf_abs,
                                                            const double *__restrict__ mean_abs,
                                                            float *__restrict__ feat) {
  const int li = blockIdx.x * 256 + threadIdx.x;
  if (li >= a.nleaves) return;
  const int A = a.leaves[li];
  const NodeRec NA = a.nodes[A];
  const int cA = node_count(NA.cnt);
  const int root = a.root_shift < 0 ? 0 : (int)(a.key[NA.first] >> a.root_shift);
  const double side = NA.region, rside = a.nodes[root].region;
  // sum of |cir| over the leaf's blobs (formed by the S2M pass)
  const double sabs = leaf_abs[A];
  const double2 a0 = a.coef[(size_t)A * a.p], a1 = a.coef[(size_t)A * a.p + 1];
  const int par = a.gparent[A >> 2];
  const int cpar = par >= 0 ? node_count(a.nodes[par].cnt) : cA;
  const double sg = a.smax > 0.0 ? a.smax : 1.0;
  const double ma = *mean_abs > 0.0 ? *mean_abs : 1.0;
  const double tiny = 1e-300;
  float f[kIn];
  f[0] = (float)(log2(rside / side) / 16.0);
  f[1] = (float)(log2(side / sg) / 8.0);
  f[2] = (float)cA / (float)leaf_cap;
  f[3] = (float)(log10(sabs / (cA * ma) + 1e-12) / 2.0);
  f[4] = (float)(sqrt(a0.x * a0.x + a0.y * a0.y) / (sabs + tiny));
  f[5] = (float)(sqrt(a1.x * a1.x + a1.y * a1.y) / (sabs * side + tiny));
  f[6] = (float)(log2(cA / ((side / sg) * (side / sg)) + 1e-3) / 8.0);
  f[7] = (float)(log2((double)cpar / cA) / 4.0);
  for (int k = 0; k < kIn; k++) feat[(size_t)li * kIn + k] = f[k];
}

// mean |cir| in a fixed order: block partials (grid-stride sums, then a
// fixed butterfly and the four waves in order), then one block adds the
// partials in block order -- the same bits on every run.
constexpr int kAbsBlocks = 1024;
static_assert(kMeanAbsWords >= 1 + kAbsBlocks, "mean |cir| scratch");
__global__ __launch_bounds__(256) void abs_sum_kernel(const double *__restrict__ v, int n,
                                                      double *__restrict__ part) {
  double s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += fabs(v[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}

// out[0] = (sum of the nb <= 1024 partials at out[1..nb]) / n, in a fixed
// order: lane l adds partials l, l + 64, ... in turn, then a butterfly across
// the wave -- the same additions in the same order on every run.
__global__ __launch_bounds__(64) void abs_mean_kernel(double *__restrict__ out, int nb, int n) {
  const int lane = threadIdx.x;
  double s = 0.0;
  for (int b = lane; b < nb; b += 64) s += out[1 + b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) out[0] = s / (double)(n > 0 ? n : 1);
}

int launch_leaf_features(const FmmArgs &a, int leaf_cap, const double *d_cir, int n,
                         const double *d_leaf_abs, double *d_mean_abs, float *d_feat,
                         hipStream_t s) {
  if (a.nleaves <= 0) return 0;
  int nb = (n + 255) / 256;
  nb = nb > kAbsBlocks ? kAbsBlocks : (nb < 1 ? 1 : nb);
  abs_sum_kernel<<<nb, 256, 0, s>>>(d_cir, n, d_mean_abs + 1);
  abs_mean_kernel<<<1, 64, 0, s>>>(d_mean_abs, nb, n);
  leaf_features_kernel<<<(a.nleaves + 255) / 256, 256, 0, s>>>(a, leaf_cap, d_leaf_abs, d_mean_abs,
                                                                d_feat);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// MLP
// ---------------------------------------------------------------------------
// One thread per sample; weights staged in LDS.  With TRAIN also the
// backward pass: per-sample activations and deltas for the weight gradients
// of mean((out - y)^2) over the n samples (training; inference runs
// mlp_infer_kernel below, whose multiply-adds are fused, so its outputs
// differ from this forward pass by rounding only).
template <bool TRAIN>
__global__ __launch_bounds__(256) void mlp_kernel(const float *__restrict__ wts, int n,
                                                  const float *__restrict__ x,
                                                  const float *__restrict__ y,
                                                  float *__restrict__ out,
                                                  float *__restrict__ h1s, float *__restrict__ h2s,
                                                  float *__restrict__ d1s, float *__restrict__ d2s,
                                                  float *__restrict__ d3s,
                                                  float *__restrict__ loss) {
  __shared__ float W[kNW];
  for (int k = threadIdx.x; k < kNW; k += 256) W[k] = wts[k];
  __syncthreads();
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  float xin[kIn];
#pragma unroll
  for (int k = 0; k < kIn; k++) xin[k] = x[(size_t)b * kIn + k];
  float h1[kH];
#pragma unroll
  for (int i = 0; i < kH; i++) {
    float s = W[oB1 + i];
#pragma unroll
    for (int k = 0; k < kIn; k++) s += W[oW1 + i * kIn + k] * xin[k];
    h1[i] = s > 0.f ? s : 0.f;
  }
  float o[kOut] = {W[oB3], W[oB3 + 1]};
  float h2[TRAIN ? kH : 1];
  for (int j = 0; j < kH; j++) {
    float s = W[oB2 + j];
#pragma unroll
    for (int i = 0; i < kH; i++) s += W[oW2 + j * kH + i] * h1[i];
    s = s > 0.f ? s : 0.f;
    if (TRAIN) h2[j] = s;
    o[0] += W[oW3 + j] * s;
    o[1] += W[oW3 + kH + j] * s;
  }
  if (out) {
    out[(size_t)b * kOut] = o[0];
    out[(size_t)b * kOut + 1] = o[1];
  }
  if (!TRAIN) return;
  const float inv_n = 1.0f / (float)n;
  float d3[kOut];
  float l = 0.f;
#pragma unroll
  for (int k = 0; k < kOut; k++) {
    const float e = o[k] - y[(size_t)b * kOut + k];
    l += e * e;
    d3[k] = 2.0f * e * inv_n;
    d3s[(size_t)b * kOut + k] = d3[k];
  }
  loss[b] = l * inv_n;
  float d2[kH];
  for (int j = 0; j < kH; j++) {
    const float g = W[oW3 + j] * d3[0] + W[oW3 + kH + j] * d3[1];
    d2[j] = h2[j] > 0.f ? g : 0.f;
    h2s[(size_t)b * kH + j] = h2[j];
    d2s[(size_t)b * kH + j] = d2[j];
  }
#pragma unroll
  for (int i = 0; i < kH; i++) {
    float g = 0.f;
    for (int j = 0; j < kH; j++) g += W[oW2 + j * kH + i] * d2[j];
    d1s[(size_t)b * kH + i] = h1[i] > 0.f ? g : 0.f;
    h1s[(size_t)b * kH + i] = h1[i];
  }
}

// gradient of weight k = sum over samples of (delta of its output) x (its input)
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(int n, const float *__restrict__ x,
                                                        const float *__restrict__ h1s,
                                                        const float *__restrict__ h2s,
                                                        const float *__restrict__ d1s,
                                                        const float *__restrict__ d2s,
                                                        const float *__restrict__ d3s,
                                                        float *__restrict__ grad) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= kNW) return;
  const float *dl;
  const float *in;
  int dst, dout, ist, iin;  // delta stride/index, input stride/index (iin < 0: bias)
  if (k < oB1) {
    dl = d1s; dst = kH; dout = (k - oW1) / kIn; in = x; ist = kIn; iin = (k - oW1) % kIn;
  } else if (k < oW2) {
    dl = d1s; dst = kH; dout = k - oB1; in = nullptr; ist = 0; iin = -1;
  } else if (k < oB2) {
    dl = d2s; dst = kH; dout = (k - oW2) / kH; in = h1s; ist = kH; iin = (k - oW2) % kH;
  } else if (k < oW3) {
    dl = d2s; dst = kH; dout = k - oB2; in = nullptr; ist = 0; iin = -1;
  } else if (k < oB3) {
    dl = d3s; dst = kOut; dout = (k - oW3) / kH; in = h2s; ist = kH; iin = (k - oW3) % kH;
  } else {
    dl = d3s; dst = kOut; dout = k - oB3; in = nullptr; ist = 0; iin = -1;
  }
  float s = 0.f;
  for (int b = 0; b < n; b++) {
    const float d = dl[(size_t)b * dst + dout];
    s += iin < 0 ? d : d * in[(size_t)b * ist + iin];
  }
  grad[k] = s;
}

__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ w, const float *__restrict__ g,
                                                   float *__restrict__ m, float *__restrict__ v,
                                                   float lr, float b1t, float b2t) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= kNW) return;
  const float b1 = 0.9f, b2 = 0.999f;
  m[k] = b1 * m[k] + (1.f - b1) * g[k];
  v[k] = b2 * v[k] + (1.f - b2) * g[k] * g[k];
  const float mh = m[k] / (1.f - b1t), vh = v[k] / (1.f - b2t);
  w[k] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// Inference: one thread per row, the weights read at wave-uniform addresses
// (scalar loads; no copy into LDS per block) and every multiply-add fused.
__global__ __launch_bounds__(256) void mlp_infer_kernel(const float *__restrict__ wts, int n,
                                                        const float *__restrict__ x,
                                                        float *__restrict__ out) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  float xin[kIn];
#pragma unroll
  for (int k = 0; k < kIn; k++) xin[k] = x[(size_t)b * kIn + k];
  float h1[kH];
#pragma unroll
  for (int i = 0; i < kH; i++) {
    float s = wts[oB1 + i];
#pragma unroll
    for (int k = 0; k < kIn; k++) s = fmaf(wts[oW1 + i * kIn + k], xin[k], s);
    h1[i] = s > 0.f ? s : 0.f;
  }
  float o0 = wts[oB3], o1 = wts[oB3 + 1];
  // four hidden units at a time: four independent multiply-add chains
  for (int j = 0; j < kH; j += 4) {
    float s[4];
#pragma unroll
    for (int u = 0; u < 4; u++) s[u] = wts[oB2 + j + u];
#pragma unroll
    for (int i = 0; i < kH; i++) {
#pragma unroll
      for (int u = 0; u < 4; u++) s[u] = fmaf(wts[oW2 + (j + u) * kH + i], h1[i], s[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const float h = s[u] > 0.f ? s[u] : 0.f;
      o0 = fmaf(wts[oW3 + j + u], h, o0);
      o1 = fmaf(wts[oW3 + kH + j + u], h, o1);
    }
  }
  out[(size_t)b * kOut] = o0;
  out[(size_t)b * kOut + 1] = o1;
}

int launch_mlp_forward(const float *d_w, int n, const float *d_x, float *d_out, hipStream_t s) {
  if (n <= 0) return 0;
  mlp_infer_kernel<<<(n + 255) / 256, 256, 0, s>>>(d_w, n, d_x, d_out);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_mlp_grad(const float *d_w, int n, const float *d_x, const float *d_y, float *d_scratch,
                    float *d_grad, float *d_loss, hipStream_t s) {
  if (n <= 0) return 0;
  float *h1 = d_scratch, *h2 = h1 + (size_t)n * kH, *d1 = h2 + (size_t)n * kH,
        *d2 = d1 + (size_t)n * kH, *d3 = d2 + (size_t)n * kH;
  mlp_kernel<true><<<(n + 255) / 256, 256, 0, s>>>(d_w, n, d_x, d_y, nullptr, h1, h2, d1, d2, d3,
                                                   d_loss);
  FMM_HIP_TRY(hipGetLastError());
  mlp_wgrad_kernel<<<(kNW + 255) / 256, 256, 0, s>>>(n, d_x, h1, h2, d1, d2, d3, d_grad);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

size_t mlp_scratch_floats(int n) { return (size_t)n * (4 * kH + kOut); }

int launch_adam(float *d_w, const float *d_g, float *d_m, float *d_v, float lr, int t,
                hipStream_t s) {
  const float b1t = powf(0.9f, (float)t), b2t = powf(0.999f, (float)t);
  adam_kernel<<<(kNW + 255) / 256, 256, 0, s>>>(d_w, d_g, d_m, d_v, lr, b1t, b2t);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// target-leaf index of every blob (caller order)
__global__ __launch_bounds__(256) void blob_leaf_kernel(const FmmArgs a, int32_t *__restrict__ out) {
  const int li = blockIdx.x * 256 + threadIdx.x;
  if (li >= a.nleaves) return;
  const NodeRec NA = a.nodes[a.leaves[li]];
  const int c = node_count(NA.cnt);
  for (int j = 0; j < c; j++) out[a.perm[NA.first + j]] = li;
}

int launch_blob_leaf(const FmmArgs &a, int32_t *d_out, hipStream_t s) {
  if (a.nleaves <= 0) return 0;
  blob_leaf_kernel<<<(a.nleaves + 255) / 256, 256, 0, s>>>(a, d_out);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// MLP outputs -> per-leaf (terms, cut); non-finite outputs fall back to the
// fixed order and radius.
__global__ __launch_bounds__(256) void leaf_params_kernel(const float *__restrict__ y, int n,
                                                          float p_margin, float cut_margin,
                                                          float cut_min, int p_max,
                                                          float cut_fallback,
                                                          float2 *__restrict__ par) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float y0 = y[2 * i], y1 = y[2 * i + 1];
  float p, c;
  if (!isfinite(y0) || !isfinite(y1)) {
    p = (float)p_max;
    c = cut_fallback;
  } else {
    p = ceilf(11.f + 5.f * y0 + p_margin);
    p = fminf(fmaxf(p, 4.f), (float)p_max);
    c = 3.f + 2.f * y1 + cut_margin;
    c = fminf(fmaxf(c, cut_min), 6.f);
  }
  par[i] = make_float2(p, c);
}

int launch_leaf_params(const float *d_y, int n, float p_margin, float cut_margin, float cut_min,
                       int p_max, float cut_fallback, float2 *d_par, hipStream_t s) {
  if (n <= 0) return 0;
  leaf_params_kernel<<<(n + 255) / 256, 256, 0, s>>>(d_y, n, p_margin, cut_margin, cut_min, p_max,
                                                     cut_fallback, d_par);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace fmm2d
