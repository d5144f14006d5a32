// step_kernels.hip -- the solver's per-step updates around the velocity
// evaluation, on device-resident blobs (gfx950), so the blobs stay in HBM
// across time steps.
//
//   drift     x1 = x0 + dt (23/12 v1 - 4/3 v2 + 5/12 v3), third-order
//             Adams-Bashforth on the velocity history (e_drift.f90:29-44),
//             optional reflection y1 -> 2 d - y1 below the mirror line
//             (e_drift.f90:46)
//   history   v3 <- v2, v2 <- v1, and on the first step v2 <- v1 first
//             (e_velo_update.f90:17-38)
//   store     v1 <- (u_inf - omz y, v_inf + omz x) + induced velocity
//             (e_velo_mpi.f90:44-47; panel and sheet terms are not part of
//             this library)
//   core      sigma <- sqrt(sigma^2 + 2.2418^2 dt / Re) (e_core.f90:15-19)
//   mirror    image blobs (x, 2 d - y), circulation -cir, point-vortex
//             kernel (core 0), as biot_b adds them (d_biot.f90:39-53)
// The velocity history is laid out like the solver's vor_vb(2, 3, n):
// hist[(i * 3 + k) * 2 + c].
#include <hip/hip_runtime.h>

#include "fmm2d_internal.h"

namespace fmm2d {

__global__ __launch_bounds__(256) void drift_ab3_kernel(int n, double2 *__restrict__ xy,
                                                        const double2 *__restrict__ hist, double dt,
                                                        int mirror, double dm) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2 v1 = hist[3 * i], v2 = hist[3 * i + 1], v3 = hist[3 * i + 2];
  const double vmx = 23.0 / 12.0 * v1.x - 4.0 / 3.0 * v2.x + 5.0 / 12.0 * v3.x;
  const double vmy = 23.0 / 12.0 * v1.y - 4.0 / 3.0 * v2.y + 5.0 / 12.0 * v3.y;
  double2 p = xy[i];
  p.x = p.x + vmx * dt;
  p.y = p.y + vmy * dt;
  if (mirror && p.y <= dm) p.y = 2.0 * dm - p.y;
  xy[i] = p;
}

__global__ __launch_bounds__(256) void history_shift_kernel(int n, double2 *__restrict__ hist,
                                                            int first) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2 v1 = hist[3 * i];
  const double2 v2 = first ? v1 : hist[3 * i + 1];
  hist[3 * i + 2] = v2;
  hist[3 * i + 1] = v1;
}

__global__ __launch_bounds__(256) void store_velocity_kernel(int n, const double2 *__restrict__ xy,
                                                             const double2 *__restrict__ uv,
                                                             double2 *__restrict__ hist,
                                                             double uinf, double vinf,
                                                             double omz) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2 p = xy[i], u = uv[i];
  const double vmx = uinf + (-omz * p.y);
  const double vmy = vinf + (omz * p.x);
  hist[3 * i] = make_double2(vmx + u.x, vmy + u.y);
}

__global__ __launch_bounds__(256) void core_spread_kernel(int n, double *__restrict__ cor,
                                                          double add) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  cor[i] = sqrt(cor[i] * cor[i] + add);
}

__global__ __launch_bounds__(256) void mirror_images_kernel(int n, const double2 *__restrict__ xy,
                                                            const double *__restrict__ cir,
                                                            double dm, double2 *__restrict__ oxy,
                                                            double *__restrict__ ocir,
                                                            double *__restrict__ ocor) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2 p = xy[i];
  oxy[i] = make_double2(p.x, 2.0 * dm - p.y);
  ocir[i] = -cir[i];
  ocor[i] = 0.0;
}

int launch_drift_ab3(int n, double *d_xy, const double *d_hist, double dt, int mirror, double dm,
                     hipStream_t s) {
  if (n <= 0) return 0;
  drift_ab3_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, reinterpret_cast<double2 *>(d_xy),
                                                   reinterpret_cast<const double2 *>(d_hist), dt,
                                                   mirror, dm);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_history_shift(int n, double *d_hist, int first, hipStream_t s) {
  if (n <= 0) return 0;
  history_shift_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, reinterpret_cast<double2 *>(d_hist),
                                                       first);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_store_velocity(int n, const double *d_xy, const double *d_uv, double *d_hist,
                          double uinf, double vinf, double omz, hipStream_t s) {
  if (n <= 0) return 0;
  store_velocity_kernel<<<(n + 255) / 256, 256, 0, s>>>(
      n, reinterpret_cast<const double2 *>(d_xy), reinterpret_cast<const double2 *>(d_uv),
      reinterpret_cast<double2 *>(d_hist), uinf, vinf, omz);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_core_spread(int n, double *d_cor, double dt, double re, hipStream_t s) {
  if (n <= 0) return 0;
  const double corc = 2.2418;
  core_spread_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, d_cor, (corc * corc) * dt / re);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

int launch_mirror_images(int n, const double *d_xy, const double *d_cir, double dm, double *d_oxy,
                         double *d_ocir, double *d_ocor, hipStream_t s) {
  if (n <= 0) return 0;
  mirror_images_kernel<<<(n + 255) / 256, 256, 0, s>>>(
      n, reinterpret_cast<const double2 *>(d_xy), d_cir, dm, reinterpret_cast<double2 *>(d_oxy),
      d_ocir, d_ocor);
  FMM_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace fmm2d
