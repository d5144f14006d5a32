// noalloc.hip -- the FP64 direct sum for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  fmm_direct_device keeps its per-chunk partial sums in a
// library-allocated workspace; this variant keeps them in `scratch`, a
// file-scope device array of fixed size that every GPU the library runs on
// holds until the process ends.  Same kernels, same source chunking, so the
// results equal fmm_direct_device's bit for bit.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fmm2d.h"
#include "fmm2d_internal.h"

namespace fmm2d {

// 2^20 partial sums (16 MiB): chunks * nt of the direct sum stays below this
// for every nt <= 2^20 (direct_chunks keeps chunks * ceil(nt / 256) near 2048).
constexpr size_t kScratchPairs = size_t(1) << 20;
static __device__ __align__(256) double2 scratch[kScratchPairs];

// Device address of `scratch` on each GPU, resolved once per device with that
// device current (the array's name in host code is only a host-side handle).
constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_address(double2 **out) {
  int dev = -1;
  FMM_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("fmm_direct_device_noalloc: device %d out of range (0..%d)", dev, kMaxDevices - 1);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    FMM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("fmm_direct_device_noalloc: no device address for the scratch array");
      return -1;
    }
    g_scratch_dev[dev] = p;
  }
  *out = static_cast<double2 *>(g_scratch_dev[dev]);
  return 0;
}

}  // namespace fmm2d

extern "C" int fmm_direct_device_noalloc(int ns, const double *d_xy, const double *d_cir,
                                         const double *d_cor, int nt, const double *d_txy,
                                         double *d_uv, int style, void *stream) {
  using namespace fmm2d;
  if (style != 0 && style != 1) {
    set_error("fmm_direct_device_noalloc: style must be 0 or 1");
    return -1;
  }
  if (ns < 0 || nt < 0) {
    set_error("fmm_direct_device_noalloc: ns=%d, nt=%d must be >= 0", ns, nt);
    return -1;
  }
  if (nt == 0) return 0;
  const size_t need = (size_t)direct_chunks(ns, nt) * (size_t)nt;
  if (need > kScratchPairs) {
    set_error("fmm_direct_device_noalloc: %d targets need %zu partial sums, more than the %zu "
              "of the fixed scratch array (at most %zu targets)",
              nt, need, kScratchPairs, kScratchPairs);
    return -1;
  }
  double2 *ws = nullptr;
  if (scratch_address(&ws)) return -1;
  return launch_direct(ns, d_xy, d_cir, d_cor, nt, d_txy, d_uv, style, ws, kScratchPairs,
                       static_cast<hipStream_t>(stream));
}
