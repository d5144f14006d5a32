// tree_kernels.hip -- quadtree construction on the GPU (gfx950).
//
//   bbox      root square from exact min/max (quadtree.c:81-99)
//   keygen    per blob, the quadrant path from the root, 2 bits per level,
//             with the parent->child centre recurrence of quadtree.c:61-79 and
//             the tie rule x >= cx, y <= cy of quadtree.c:29-35
//   radix     stable LSD sort of (key, blob id), 8-bit digits, LDS ranking
//   levels    two launches per level with the counts kept on the device:
//             children of every split node by wave searches on the sorted
//             keys, leaf iff count < num_p (quadtree.c:197-245); the split
//             children compacted into the next level's list by a single-pass
//             scan with decoupled look-back
//   leaf      within every leaf, blobs in ascending id (the stable partition
//             order of quadtree.c:219-225)
#include <hip/hip_runtime.h>

#include "fmm2d_internal.h"

namespace fmm2d {

// ---------------------------------------------------------------------------
// wave / block helpers (wave64)
// ---------------------------------------------------------------------------
__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread across a 256-thread block.
__device__ inline uint32_t block_excl_scan_256(uint32_t v, uint32_t *lds4, uint32_t *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) lds4[w] = inc;
  __syncthreads();
  uint32_t off = 0;
  for (int k = 0; k < w; k++) off += lds4[k];
  if (total) *total = lds4[0] + lds4[1] + lds4[2] + lds4[3];
  __syncthreads();
  return off + inc - v;
}

// ---------------------------------------------------------------------------
// exclusive scan of u32 (1024 values per block, recursive over block sums)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void scan_block_kernel(const uint32_t *__restrict__ in,
                                                         uint32_t *__restrict__ out,
                                                         uint32_t *__restrict__ sums, int n) {
  __shared__ uint32_t lds4[4];
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t a[4];
#pragma unroll
  for (int j = 0; j < 4; j++) a[j] = (base + j < n) ? in[base + j] : 0u;
  uint32_t t = a[0] + a[1] + a[2] + a[3];
  uint32_t tot;
  uint32_t ex = block_excl_scan_256(t, lds4, &tot);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (base + j < n) out[base + j] = ex;
    ex += a[j];
  }
  if (threadIdx.x == 0 && sums) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void scan_add_kernel(uint32_t *__restrict__ out,
                                                       const uint32_t *__restrict__ offs, int n) {
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  const uint32_t o = offs[blockIdx.x];
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (base + j < n) out[base + j] += o;
}

// Workspace: the scan needs block-sum arrays for each recursion level; they
// live in a grow-only device buffer owned by this file.
static uint32_t *g_scan_ws = nullptr;
static size_t g_scan_ws_words = 0;

static int scan_rec(const uint32_t *d_in, uint32_t *d_out, int n, uint32_t *d_total, uint32_t *ws,
                    hipStream_t s) {
  const int nb = (n + 1023) / 1024;
  if (nb <= 1) {
    scan_block_kernel<<<1, 256, 0, s>>>(d_in, d_out, d_total, n);
    FMM_HIP_TRY(hipGetLastError());
    return 0;
  }
  uint32_t *sums = ws, *offs = ws + nb;
  scan_block_kernel<<<nb, 256, 0, s>>>(d_in, d_out, sums, n);
  FMM_HIP_TRY(hipGetLastError());
  if (scan_rec(sums, offs, nb, d_total, ws + 2 * nb, s) != 0) return -1;
  scan_add_kernel<<<nb, 256, 0, s>>>(d_out, offs, n);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// One workspace serves every scan: scans on two streams must be ordered
// (the callers order them by events).
int scan_u32(const uint32_t *d_in, uint32_t *d_out, int n, uint32_t *d_total, hipStream_t s) {
  if (n <= 0) {
    if (d_total) FMM_HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(uint32_t), s));
    return 0;
  }
  size_t need = 64;
  for (size_t m = (size_t)n; m > 1; m = (m + 1023) / 1024) need += 2 * ((m + 1023) / 1024) + 2;
  if (need > g_scan_ws_words) {
    if (g_scan_ws) {
      FMM_HIP_TRY(hipStreamSynchronize(s));
      FMM_HIP_TRY(hipFree(g_scan_ws));
      g_scan_ws = nullptr;
    }
    size_t words = need * 2;
    FMM_HIP_TRY(hipMalloc(&g_scan_ws, words * sizeof(uint32_t)));
    g_scan_ws_words = words;
  }
  return scan_rec(d_in, d_out, n, d_total, g_scan_ws, s);
}

// ---------------------------------------------------------------------------
// bbox: exact min / max of x and y, then the root square
// ---------------------------------------------------------------------------
constexpr int kBboxBlocks = 512;

__global__ __launch_bounds__(256) void bbox_partial_kernel(const double2 *__restrict__ xy, int n,
                                                           double *__restrict__ part) {
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double2 p = xy[i];
    mnx = fmin(mnx, p.x);
    mxx = fmax(mxx, p.x);
    mny = fmin(mny, p.y);
    mxy = fmax(mxy, p.y);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mnx = fmin(mnx, __shfl_xor(mnx, o, 64));
    mxx = fmax(mxx, __shfl_xor(mxx, o, 64));
    mny = fmin(mny, __shfl_xor(mny, o, 64));
    mxy = fmax(mxy, __shfl_xor(mxy, o, 64));
  }
  __shared__ double s[4][4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s[w][0] = mnx; s[w][1] = mxx; s[w][2] = mny; s[w][3] = mxy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) {
      s[0][0] = fmin(s[0][0], s[k][0]);
      s[0][1] = fmax(s[0][1], s[k][1]);
      s[0][2] = fmin(s[0][2], s[k][2]);
      s[0][3] = fmax(s[0][3], s[k][3]);
    }
    for (int k = 0; k < 4; k++) part[4 * blockIdx.x + k] = s[0][k];
  }
}

__global__ __launch_bounds__(64) void bbox_final_kernel(const double *__restrict__ part, int nb,
                                                        RootBox *__restrict__ root) {
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
  for (int b = threadIdx.x; b < nb; b += 64) {
    mnx = fmin(mnx, part[4 * b + 0]);
    mxx = fmax(mxx, part[4 * b + 1]);
    mny = fmin(mny, part[4 * b + 2]);
    mxy = fmax(mxy, part[4 * b + 3]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mnx = fmin(mnx, __shfl_xor(mnx, o, 64));
    mxx = fmax(mxx, __shfl_xor(mxx, o, 64));
    mny = fmin(mny, __shfl_xor(mny, o, 64));
    mxy = fmax(mxy, __shfl_xor(mxy, o, 64));
  }
  if (threadIdx.x == 0) {
    // quadtree.c:87-98: side = max(|dx| + 1e-4, |dy| + 1e-4), x side wins only if strictly larger
    double rx = fabs(mxx - mnx) + 0.0001;
    double ry = fabs(mxy - mny) + 0.0001;
    root->region = rx > ry ? rx : ry;
    root->cx = 0.5 * (mxx + mnx);
    root->cy = 0.5 * (mxy + mny);
    root->pad = 0.0;
  }
}

int launch_bbox(const double *d_xy, int n, RootBox *d_root, double *d_partial, hipStream_t s) {
  int nb = (n + 255) / 256;
  if (nb > kBboxBlocks) nb = kBboxBlocks;
  if (nb < 1) nb = 1;
  bbox_partial_kernel<<<nb, 256, 0, s>>>(reinterpret_cast<const double2 *>(d_xy), n, d_partial);
  FMM_HIP_TRY(hipGetLastError());
  bbox_final_kernel<<<1, 64, 0, s>>>(d_partial, nb, d_root);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// keygen: quadrant path of each blob, most significant digit = level 1
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void keygen_kernel(const double2 *__restrict__ xy, int n,
                                                     const RootBox *__restrict__ root, int depth,
                                                     uint64_t *__restrict__ key,
                                                     uint32_t *__restrict__ idx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2 p = xy[i];
  double cx = root->cx, cy = root->cy, reg = root->region;
  uint64_t k = 0;
  for (int l = 0; l < depth; l++) {
    const bool right = p.x >= cx;
    const bool down = p.y <= cy;
    k = (k << 2) | (uint64_t)((right ? 1 : 0) + (down ? 2 : 0));
    const double h = 0.25 * reg;
    cx = right ? cx + h : cx - h;
    cy = down ? cy - h : cy + h;
    reg = 0.5 * reg;
  }
  key[i] = k;
  idx[i] = (uint32_t)i;
}

int launch_keygen(const double *d_xy, int n, const RootBox *d_root, int depth, uint64_t *d_key,
                  uint32_t *d_idx, hipStream_t s) {
  keygen_kernel<<<(n + 255) / 256, 256, 0, s>>>(reinterpret_cast<const double2 *>(d_xy), n, d_root,
                                                 depth, d_key, d_idx);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// forests: many independent systems, concatenated; system s owns blobs
// [off[s], off[s+1]).  One block per system for its root square; the key of a
// blob is (s << 2D) | quadrant path, so one sort orders systems, then paths.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bbox_segments_kernel(const double2 *__restrict__ xy,
                                                            const int32_t *__restrict__ off,
                                                            RootBox *__restrict__ roots) {
  const int s = blockIdx.x;
  const int b = off[s], e = off[s + 1];
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
  for (int i = b + threadIdx.x; i < e; i += 256) {
    const double2 p = xy[i];
    mnx = fmin(mnx, p.x);
    mxx = fmax(mxx, p.x);
    mny = fmin(mny, p.y);
    mxy = fmax(mxy, p.y);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mnx = fmin(mnx, __shfl_xor(mnx, o, 64));
    mxx = fmax(mxx, __shfl_xor(mxx, o, 64));
    mny = fmin(mny, __shfl_xor(mny, o, 64));
    mxy = fmax(mxy, __shfl_xor(mxy, o, 64));
  }
  __shared__ double sm[4][4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sm[w][0] = mnx; sm[w][1] = mxx; sm[w][2] = mny; sm[w][3] = mxy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) {
      mnx = fmin(mnx, sm[k][0]);
      mxx = fmax(mxx, sm[k][1]);
      mny = fmin(mny, sm[k][2]);
      mxy = fmax(mxy, sm[k][3]);
    }
    mnx = fmin(mnx, sm[0][0]); mxx = fmax(mxx, sm[0][1]);
    mny = fmin(mny, sm[0][2]); mxy = fmax(mxy, sm[0][3]);
    RootBox r;
    if (e > b) {
      const double rx = fabs(mxx - mnx) + 0.0001;
      const double ry = fabs(mxy - mny) + 0.0001;
      r.region = rx > ry ? rx : ry;
      r.cx = 0.5 * (mxx + mnx);
      r.cy = 0.5 * (mxy + mny);
    } else {
      r.region = 0.0; r.cx = 0.0; r.cy = 0.0;
    }
    r.pad = 0.0;
    roots[s] = r;
  }
}

__global__ __launch_bounds__(256) void keygen_forest_kernel(const double2 *__restrict__ xy, int n,
                                                            const int32_t *__restrict__ off,
                                                            int nsys,
                                                            const RootBox *__restrict__ roots,
                                                            int depth, uint64_t *__restrict__ key,
                                                            uint32_t *__restrict__ idx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  // system of blob i: last s with off[s] <= i
  int lo = 0, hi = nsys;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  const RootBox R = roots[lo];
  const double2 p = xy[i];
  double cx = R.cx, cy = R.cy, reg = R.region;
  uint64_t k = 0;
  for (int l = 0; l < depth; l++) {
    const bool right = p.x >= cx;
    const bool down = p.y <= cy;
    k = (k << 2) | (uint64_t)((right ? 1 : 0) + (down ? 2 : 0));
    const double h = 0.25 * reg;
    cx = right ? cx + h : cx - h;
    cy = down ? cy - h : cy + h;
    reg = 0.5 * reg;
  }
  key[i] = ((uint64_t)lo << (2 * depth)) | k;
  idx[i] = (uint32_t)i;
}

// Roots 0..nsys-1 (padding nodes up to nroot_slots, a multiple of 4).
__global__ __launch_bounds__(256) void init_forest_kernel(NodeRec *__restrict__ nodes,
                                                          int32_t *__restrict__ gparent,
                                                          const RootBox *__restrict__ roots,
                                                          const int32_t *__restrict__ off,
                                                          int nsys, int nroot_slots) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nroot_slots) return;
  NodeRec r;
  if (s < nsys) {
    const RootBox b = roots[s];
    const int cnt = off[s + 1] - off[s];
    r.cx = b.cx; r.cy = b.cy; r.region = b.region;
    r.first = off[s];
    r.cnt = cnt > 0 ? (cnt | kInternalBit) : 0;
  } else {
    r.cx = 0.0; r.cy = 0.0; r.region = 0.0; r.first = 0; r.cnt = 0;
  }
  nodes[s] = r;
  if ((s & 3) == 0) gparent[s >> 2] = -1;
}

int launch_forest_roots(const double *d_xy, int n, const int32_t *d_off, int nsys, int depth,
                        RootBox *d_roots, uint64_t *d_key, uint32_t *d_idx, NodeRec *d_nodes,
                        int32_t *d_gparent, int nroot_slots, hipStream_t s) {
  bbox_segments_kernel<<<nsys, 256, 0, s>>>(reinterpret_cast<const double2 *>(d_xy), d_off,
                                            d_roots);
  FMM_HIP_TRY(hipGetLastError());
  keygen_forest_kernel<<<(n + 255) / 256, 256, 0, s>>>(reinterpret_cast<const double2 *>(d_xy), n,
                                                        d_off, nsys, d_roots, depth, d_key, d_idx);
  FMM_HIP_TRY(hipGetLastError());
  init_forest_kernel<<<(nroot_slots + 255) / 256, 256, 0, s>>>(d_nodes, d_gparent, d_roots, d_off,
                                                               nsys, nroot_slots);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  A tile is 4096 pairs, 1024 per wave: item j
// of lane l of wave w is pair w*1024 + j*64 + l, so the tile order that must
// be preserved is (wave, item, lane).  Lanes holding the same digit are
// found with eight ballots; the leader advances a per-wave digit counter in
// LDS.  Pairs are staged in LDS in tile-sorted order and written out in runs.
// ---------------------------------------------------------------------------
constexpr int kRsThreads = 256, kRsItems = 16, kRsTile = kRsThreads * kRsItems;

__device__ inline uint64_t peers_of(uint32_t d, uint64_t valid) {
  uint64_t m = valid;
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

__global__ __launch_bounds__(256) void rs_hist_kernel(const uint64_t *__restrict__ key, int n,
                                                      int shift, uint32_t *__restrict__ hist,
                                                      int ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wbase = blockIdx.x * kRsTile + w * (kRsItems * 64);
  for (int j = 0; j < kRsItems; j++) {
    const int i = wbase + j * 64 + lane;
    const bool ok = i < n;
    const uint32_t d = ok ? (uint32_t)(key[i] >> shift) & 255u : 0u;
    const uint64_t valid = __ballot(ok);
    if (ok) {
      const uint64_t pm = peers_of(d, valid);
      if (lane == __ffsll((unsigned long long)pm) - 1) atomicAdd(&h[d], (uint32_t)__popcll(pm));
    }
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(256) void rs_scatter_kernel(const uint64_t *__restrict__ kin,
                                                         const uint32_t *__restrict__ vin,
                                                         uint64_t *__restrict__ kout,
                                                         uint32_t *__restrict__ vout, int n,
                                                         int shift,
                                                         const uint32_t *__restrict__ hscan,
                                                         int ntiles) {
  __shared__ uint32_t wcnt[4][256];
  __shared__ uint32_t dstart[256];
  __shared__ uint32_t gbase[256];
  __shared__ uint32_t lds4[4];
  __shared__ uint64_t skey[kRsTile];
  __shared__ uint32_t sval[kRsTile];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int k = tid; k < 4 * 256; k += 256) (&wcnt[0][0])[k] = 0;
  __syncthreads();
  const int tbase = blockIdx.x * kRsTile;
  const int wbase = tbase + w * (kRsItems * 64);
  const uint64_t lt = (1ull << lane) - 1ull;
  uint64_t k[kRsItems];
  uint32_t v[kRsItems], r[kRsItems];
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    const int i = wbase + j * 64 + lane;
    const bool ok = i < n;
    k[j] = ok ? kin[i] : 0ull;
    v[j] = ok ? vin[i] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    const int i = wbase + j * 64 + lane;
    const bool ok = i < n;
    const uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
    const uint64_t valid = __ballot(ok);
    r[j] = 0;
    if (ok) {
      const uint64_t pm = peers_of(d, valid);
      const uint32_t prior = wcnt[w][d];
      r[j] = prior + (uint32_t)__popcll(pm & lt);
      if (lane == __ffsll((unsigned long long)pm) - 1) wcnt[w][d] = prior + (uint32_t)__popcll(pm);
    }
  }
  __syncthreads();
  // per digit: exclusive prefix over waves, then over digits
  uint32_t tot = 0;
  {
    const int d = tid;
    uint32_t run = 0;
    for (int ww = 0; ww < 4; ww++) {
      const uint32_t t = wcnt[ww][d];
      wcnt[ww][d] = run;
      run += t;
    }
    tot = run;
  }
  const uint32_t ds = block_excl_scan_256(tot, lds4, nullptr);
  dstart[tid] = ds;
  gbase[tid] = hscan[tid * ntiles + blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    const int i = wbase + j * 64 + lane;
    if (i < n) {
      const uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
      const uint32_t pos = dstart[d] + wcnt[w][d] + r[j];
      skey[pos] = k[j];
      sval[pos] = v[j];
    }
  }
  __syncthreads();
  const int cnt = min(kRsTile, n - tbase);
  for (int e = tid; e < cnt; e += 256) {
    const uint64_t kk = skey[e];
    const uint32_t d = (uint32_t)(kk >> shift) & 255u;
    const uint32_t dst = gbase[d] + (uint32_t)e - dstart[d];
    kout[dst] = kk;
    vout[dst] = sval[e];
  }
}

size_t radix_sort_hist_words(int n) {
  const size_t ntiles = (size_t)((n + kRsTile - 1) / kRsTile);
  return 256 * (ntiles > 0 ? ntiles : 1);
}

int radix_sort_pairs(uint64_t **key, uint32_t **idx, uint64_t **key_alt, uint32_t **idx_alt, int n,
                     int lo_bit, int bits, uint32_t *d_hist, uint32_t *d_hist_scan,
                     hipStream_t s) {
  if (n <= 1) return 0;
  const int ntiles = (n + kRsTile - 1) / kRsTile;
  for (int shift = lo_bit; shift < bits; shift += 8) {
    rs_hist_kernel<<<ntiles, 256, 0, s>>>(*key, n, shift, d_hist, ntiles);
    FMM_HIP_TRY(hipGetLastError());
    if (scan_u32(d_hist, d_hist_scan, 256 * ntiles, nullptr, s) != 0) return -1;
    rs_scatter_kernel<<<ntiles, 256, 0, s>>>(*key, *idx, *key_alt, *idx_alt, n, shift, d_hist_scan,
                                             ntiles);
    FMM_HIP_TRY(hipGetLastError());
    uint64_t *tk = *key; *key = *key_alt; *key_alt = tk;
    uint32_t *ti = *idx; *idx = *idx_alt; *idx_alt = ti;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// level-by-level tree build
// ---------------------------------------------------------------------------
__global__ void init_root_kernel(NodeRec *__restrict__ nodes, int32_t *__restrict__ gparent,
                                 const RootBox *__restrict__ root, int n,
                                 int32_t *__restrict__ split) {
  const int t = threadIdx.x;
  if (t < 4) {
    NodeRec r;
    if (t == 0) {
      r.cx = root->cx; r.cy = root->cy; r.region = root->region;
      r.first = 0;  // the range start; replaced by the child index when level 1 is built
      r.cnt = n | kInternalBit;
    } else {
      r.cx = 0.0; r.cy = 0.0; r.region = 0.0; r.first = 0; r.cnt = 0;
    }
    nodes[t] = r;
  }
  if (t == 0) {
    gparent[0] = -1;
    split[0] = 0;
  }
}

int launch_init_root(NodeRec *d_nodes, int32_t *d_gparent, const RootBox *d_root, int n,
                     int32_t *d_split, hipStream_t s) {
  init_root_kernel<<<1, 64, 0, s>>>(d_nodes, d_gparent, d_root, n, d_split);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// First index in [lo, hi) whose digit at `shift` is >= q (digits do not
// decrease there), found by one wave: each step probes 64 evenly spaced keys
// and keeps the gap where the digit first reaches q (about four dependent
// loads for a million keys instead of twenty).
__device__ inline int wave_lower_bound(const uint64_t *__restrict__ key, int lo, int hi, int shift,
                                       uint32_t q, int lane) {
  while (hi - lo > 64) {
    const int pos = lo + (int)(((int64_t)(hi - lo) * (lane + 1)) / 65);
    const bool ge = (uint32_t)((key[pos] >> shift) & 3u) >= q;
    const uint64_t m = __builtin_amdgcn_ballot_w64(ge);
    if (m == 0) {
      lo = __builtin_amdgcn_readlane(pos, 63) + 1;
    } else {
      const int f = __builtin_ctzll(m);
      const int pf = __builtin_amdgcn_readlane(pos, f);
      if (f > 0) lo = __builtin_amdgcn_readlane(pos, f - 1) + 1;
      hi = pf;
    }
  }
  const int i = lo + lane;
  const uint64_t m =
      __builtin_amdgcn_ballot_w64(i < hi && (uint32_t)((key[i] >> shift) & 3u) >= q);
  return m ? lo + __builtin_ctzll(m) : hi;
}

// Two launches per level, no host round trip: the counts of a level stay on
// the device (the level table `lt`, see fmm2d_internal.h) and the next launch
// reads them.
//   children: one wave per split node of level l-1 (grid-stride over the
//             device count): its three digit boundaries by wave searches, the
//             four children's centres by quadtree.c:61-79, leaf iff count <
//             num_p (quadtree.c:197-245), and one split flag per child;
//   compact:  the split children, in Morton order, into the next split list:
//             tiles of 4096 flags, a single-pass scan with decoupled
//             look-back (each tile publishes its count, flag A, then its
//             inclusive prefix, flag P, once the counts of its predecessors
//             back to the first P are summed; one wave reads 64 predecessors
//             per step).  A tile's predecessors hold earlier tickets of
//             running workgroups, so every wait ends; it is also bounded
//             (kLookbackSpins), past which the error flag is set and the host
//             reports a failed build.
// The range start of an internal node is kept in `first` until its own
// children are built (then `first` becomes the child index).
constexpr uint32_t kStatAgg = 1u << 30, kStatPre = 2u << 30, kStatVal = (1u << 30) - 1u;
constexpr int64_t kLookbackSpins = int64_t(1) << 24;
constexpr int kCompactTile = 4096;  // flags per compaction tile (256 threads x 16)

__device__ inline uint32_t ld_acquire(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_release(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void tree_children_kernel(
    const uint64_t *__restrict__ key, int shift, int level, int max_level, int num_p,
    NodeRec *__restrict__ nodes, int32_t *__restrict__ gparent, const int32_t *__restrict__ split,
    int node_cap, int split_cap, int32_t *__restrict__ lt, uint8_t *__restrict__ cflag) {
  const int lane = threadIdx.x & 63;
  if (lt[kLtErr] != 0) return;
  const int Sp = lt[kLtS + level - 1];
  if (Sp == 0) return;
  const int cb = lt[kLtCb + level], offn = lt[kLtOff + level], offp = lt[kLtOff + level - 1];
  const bool first_thread = blockIdx.x == 0 && threadIdx.x == 0;
  if (level > max_level) {
    if (first_thread) atomicOr(&lt[kLtErr], kLtErrDepth);
    return;
  }
  if ((int64_t)cb + 4 * (int64_t)Sp > node_cap || (int64_t)offn + 4 * (int64_t)Sp > split_cap) {
    if (first_thread) atomicOr(&lt[kLtErr], kLtErrOverflow);
    return;
  }
  for (int j = blockIdx.x * 4 + (threadIdx.x >> 6); j < Sp; j += gridDim.x * 4) {
    const int par = split[offp + j];
    const NodeRec P = nodes[par];
    const int lo = P.first, hi = lo + node_count(P.cnt);
    const int b2 = wave_lower_bound(key, lo, hi, shift, 2u, lane);
    const int b1 = wave_lower_bound(key, lo, b2, shift, 1u, lane);
    const int b3 = wave_lower_bound(key, b2, hi, shift, 3u, lane);
    const int c0 = cb + 4 * j;
    if (lane < 4) {
      const int q = lane;
      const int bq = q == 0 ? lo : q == 1 ? b1 : q == 2 ? b2 : b3;
      const int be = q == 0 ? b1 : q == 1 ? b2 : q == 2 ? b3 : hi;
      const double h = 0.25 * P.region;
      NodeRec c;
      // quadtree.c:61-79
      c.cx = (q & 1) ? P.cx + h : P.cx - h;
      c.cy = (q < 2) ? P.cy + h : P.cy - h;
      c.region = 0.5 * P.region;
      const int cnt = be - bq;
      c.first = bq;
      const bool sp = cnt >= num_p;
      c.cnt = sp ? (cnt | kInternalBit) : cnt;
      nodes[c0 + q] = c;
      cflag[4 * j + q] = sp ? 1 : 0;
    }
    if (lane == 0) {
      gparent[c0 >> 2] = par;
      nodes[par].first = c0;
    }
  }
}

__global__ __launch_bounds__(256) void tree_compact_kernel(int level, int max_level,
                                                           int32_t *__restrict__ split,
                                                           int32_t *__restrict__ lt,
                                                           const uint8_t *__restrict__ cflag,
                                                           uint32_t *__restrict__ stat,
                                                           uint32_t *__restrict__ stat_next,
                                                           int stat_len) {
  __shared__ uint32_t lds4[4];
  __shared__ int s_ticket;
  __shared__ uint32_t s_prefix;
  const int lane = threadIdx.x & 63;
  // the next level's status words start at zero (the previous level's
  // compaction, the last user of that array, has finished)
  for (int i = blockIdx.x * 256 + threadIdx.x; i < stat_len; i += gridDim.x * 256) stat_next[i] = 0u;
  if (lt[kLtErr] != 0) return;
  const int Sp = lt[kLtS + level - 1];
  const int cb = lt[kLtCb + level], offn = lt[kLtOff + level];
  const int nflag = 4 * Sp;
  const int ntiles = (nflag + kCompactTile - 1) / kCompactTile;
  if (Sp == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      lt[kLtS + level] = 0;
      lt[kLtOff + level + 1] = offn;
      lt[kLtCb + level + 1] = cb;
    }
    return;
  }
  if (ntiles > stat_len) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&lt[kLtErr], kLtErrOverflow);
    return;
  }
  for (;;) {
    if (threadIdx.x == 0) s_ticket = atomicAdd(&lt[kLtTicket + level], 1);
    __syncthreads();
    const int t = s_ticket;
    __syncthreads();
    if (t >= ntiles) break;
    const int base = t * kCompactTile + threadIdx.x * 16;
    uint32_t f[16];
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      f[k] = (base + k < nflag) ? cflag[base + k] : 0u;
      mine += f[k];
    }
    uint32_t agg = 0;
    const uint32_t ex = block_excl_scan_256(mine, lds4, &agg);
    if (threadIdx.x < 64) {
      // wave 0: publish, then look back 64 tiles per step
      uint32_t prefix = 0;
      if (t == 0) {
        if (lane == 0) st_release(&stat[0], kStatPre | agg);
      } else {
        if (lane == 0) st_release(&stat[t], kStatAgg | agg);
        int k = t - 1;
        int64_t spins = 0;
        for (;;) {
          const int idx = k - lane;
          const uint32_t v = idx >= 0 ? ld_acquire(&stat[idx]) : kStatPre;
          const uint32_t fl = v & ~kStatVal;
          const uint64_t mp = __builtin_amdgcn_ballot_w64(fl == kStatPre);
          const int fp = mp ? __builtin_ctzll(mp) : 64;  // nearest P
          const uint64_t upto = fp >= 63 ? ~0ull : ((2ull << fp) - 1ull);
          const uint64_t mx = __builtin_amdgcn_ballot_w64(fl == 0u) & upto;
          if (mx) {
            if (++spins > kLookbackSpins) {
              if (lane == 0) atomicOr(&lt[kLtErr], kLtErrLookback);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint32_t val = (lane <= fp) ? (v & kStatVal) : 0u;
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o, 64);
          prefix += val;
          if (fp < 64) break;
          k -= 64;
        }
        if (lane == 0) st_release(&stat[t], kStatPre | ((prefix + agg) & kStatVal));
      }
      if (lane == 0) {
        s_prefix = prefix;
        if (t == ntiles - 1) {
          const int total = (int)(prefix + agg);
          lt[kLtS + level] = total;
          lt[kLtOff + level + 1] = offn + total;
          lt[kLtCb + level + 1] = cb + 4 * Sp;
          if (total > 0 && level == max_level) atomicOr(&lt[kLtErr], kLtErrDepth);
        }
      }
    }
    __syncthreads();
    uint32_t pos = s_prefix + ex;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      if (f[k]) split[offn + (int)pos] = cb + base + k;
      pos += f[k];
    }
    __syncthreads();
  }
}

// lt: level 0 holds the nroots split roots (their list is already at
// split[0..nroots)), level 1's children start at nroot_slots; tickets, error
// flag and both status arrays cleared.
__global__ __launch_bounds__(256) void tree_levels_init_kernel(int32_t *__restrict__ lt, int nroots,
                                                               int nroot_slots,
                                                               uint32_t *__restrict__ stat,
                                                               int stat_words) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < stat_words; i += gridDim.x * 256) stat[i] = 0u;
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < kLtWords; i += 256) lt[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      lt[kLtS + 0] = nroots;
      lt[kLtOff + 0] = 0;
      lt[kLtOff + 1] = nroots;
      lt[kLtCb + 1] = nroot_slots;
    }
  }
}

int launch_tree_levels_init(int32_t *d_lt, int nroots, int nroot_slots, uint32_t *d_stat,
                            int stat_len, hipStream_t s) {
  const int words = 2 * stat_len;
  int nb = (words + 255) / 256;
  nb = nb < 1 ? 1 : (nb > 1024 ? 1024 : nb);
  tree_levels_init_kernel<<<nb, 256, 0, s>>>(d_lt, nroots, nroot_slots, d_stat, words);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_tree_level(const uint64_t *d_key, int depth_bits, int level, int max_level, int num_p,
                      int split_bound, NodeRec *d_nodes, int32_t *d_gparent, int32_t *d_split,
                      int node_cap, int split_cap, int32_t *d_lt, uint8_t *d_cflag,
                      uint32_t *d_stat, int stat_len, hipStream_t s) {
  const int shift = depth_bits - 2 * level;
  // one wave per split node of level - 1 (grid-stride past 2048 workgroups)
  int nb = (split_bound + 3) / 4;
  nb = nb < 1 ? 1 : (nb > 2048 ? 2048 : nb);
  tree_children_kernel<<<nb, 256, 0, s>>>(d_key, shift, level, max_level, num_p, d_nodes, d_gparent,
                                          d_split, node_cap, split_cap, d_lt, d_cflag);
  FMM_HIP_TRY(hipGetLastError());
  uint32_t *cur = d_stat + (size_t)(level & 1) * stat_len;
  uint32_t *next = d_stat + (size_t)((level + 1) & 1) * stat_len;
  int nt = (4 * split_bound + kCompactTile - 1) / kCompactTile;
  nt = nt < 1 ? 1 : (nt > 256 ? 256 : nt);
  tree_compact_kernel<<<nt, 256, 0, s>>>(level, max_level, d_split, d_lt, d_cflag, cur, next,
                                         stat_len);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int tree_level_tiles(int nchild) { return (nchild + kCompactTile - 1) / kCompactTile; }

// One wave per node: a leaf's blob ids are reordered ascending by ranking.
__global__ __launch_bounds__(256) void leaf_order_kernel(const NodeRec *__restrict__ nodes,
                                                         int nnodes,
                                                         const uint32_t *__restrict__ idx,
                                                         uint32_t *__restrict__ perm) {
  const int node = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (node >= nnodes || node == 0) return;
  const NodeRec r = nodes[node];
  if (node_internal(r.cnt)) return;
  const int c = node_count(r.cnt), lo = r.first;
  if (c <= 64) {
    // one id per lane; ranks from the other lanes' ids read with v_readlane
    const uint32_t my = lane < c ? idx[lo + lane] : 0xffffffffu;
    int rank = 0;
    for (int f = 0; f < c; f++)
      rank += (uint32_t)__builtin_amdgcn_readlane((int)my, f) < my ? 1 : 0;
    if (lane < c) perm[lo + rank] = my;
    return;
  }
  for (int e = lane; e < c; e += 64) {
    const uint32_t my = idx[lo + e];
    int rank = 0;
    for (int f = 0; f < c; f++) rank += idx[lo + f] < my ? 1 : 0;
    perm[lo + rank] = my;
  }
}

int launch_leaf_order(const NodeRec *d_nodes, int nnodes, const uint32_t *d_idx_sorted,
                      uint32_t *d_perm, int num_p, hipStream_t s) {
  (void)num_p;
  leaf_order_kernel<<<(nnodes + 3) / 4, 256, 0, s>>>(d_nodes, nnodes, d_idx_sorted, d_perm);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

// (x, y, cir, cor) of every blob in one 32-byte record (caller order), so
// the gather into leaf order reads one record per blob instead of three
// scattered words.
__global__ __launch_bounds__(256) void pack_kernel(const double2 *__restrict__ xy,
                                                   const double *__restrict__ cir,
                                                   const double *__restrict__ cor, int n,
                                                   double4 *__restrict__ rec) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2 q = xy[i];
  rec[i] = make_double4(q.x, q.y, cir[i], cor[i]);
}

__global__ __launch_bounds__(256) void gather_kernel(const uint32_t *__restrict__ perm, int n,
                                                     const double4 *__restrict__ rec,
                                                     double *__restrict__ sx, double *__restrict__ sy,
                                                     double *__restrict__ sg,
                                                     double *__restrict__ ss) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double4 r = rec[perm[i]];
  sx[i] = r.x;
  sy[i] = r.y;
  sg[i] = r.z;
  ss[i] = r.w;
}

int launch_gather(const uint32_t *d_perm, int n, const double *d_xy, const double *d_cir,
                  const double *d_cor, double4 *d_rec, double *sx, double *sy, double *sg,
                  double *ss, hipStream_t s) {
  pack_kernel<<<(n + 255) / 256, 256, 0, s>>>(reinterpret_cast<const double2 *>(d_xy), d_cir,
                                              d_cor, n, d_rec);
  FMM_HIP_TRY(hipGetLastError());
  gather_kernel<<<(n + 255) / 256, 256, 0, s>>>(d_perm, n, d_rec, sx, sy, sg, ss);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace fmm2d
