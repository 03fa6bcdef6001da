// fw_track.hip -- device trackers that read the pre-step potential of a step:
// local activation times (every crossing of every node) and spiral-wave tips.
// Built FP64 without contraction: the tip position is the reference's
// arithmetic in its order.
#include <math.h>

#include "fw_internal.hpp"
#include "fw_kernels.hpp"

namespace fw {

__global__ void k_zero_i32(int32_t* p, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = 0;
}

__global__ void k_zero_f64(double* p, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = 0.0;
}

// LocalActivationTime2DTracker.cross_threshold / _track
// (cpuwave2D/tracker/local_activation_time_2d_tracker.py:60-102), per node:
// a node crosses when u >= thr and it is not activated; it is deactivated
// when u < thr.  state: bit 0 activated, bits 1.. number of crossings so far;
// steps[c * cap + k] = step of the k-th crossing.  The layer structure of the
// reference (a new layer when a crossing node already has a time in the last
// one) depends only on these crossing steps and is rebuilt on the host.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_lat(const U* __restrict__ u,
                                                int32_t* __restrict__ state,
                                                int32_t* __restrict__ steps, int32_t cap, U thr,
                                                int32_t step, int64_t n, int32_t* overflow) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= n) return;
  const U v = u[c];
  const int32_t s = state[c];
  const int32_t activated = s & 1;
  if (v >= thr && !activated) {
    const int32_t cnt = s >> 1;
    if (cnt < cap) steps[c * cap + cnt] = step;
    else atomicOr(overflow, 1);
    state[c] = ((cnt + 1) << 1) | 1;
  } else if (v < thr && activated) {
    state[c] = s & ~1;
  }
}

// _apply_threshold (spiral_wave_core_2d_tracker.py:175-208) on the slice
// [., ., k]; si, sj: strides of i and j
template <typename U>
__device__ __forceinline__ int tip_cross(const U* __restrict__ u, int64_t c, int64_t si,
                                         int64_t sj, double thr) {
  const double a = (double)u[c], b = (double)u[c + si], d = (double)u[c + sj],
               e = (double)u[c + si + sj];
  if (a >= thr && (b < thr || d < thr || e < thr)) return 1;
  if (a < thr && (b >= thr || d >= thr || e >= thr)) return 1;
  return 0;
}

// _correct_tip_pos (spiral_wave_core_2d_tracker.py:108-172): the crossing of
// the two threshold isolines inside the cell (i, j)-(i+1, j+1)
template <typename U>
__device__ __forceinline__ bool tip_pos(const U* __restrict__ u, const U* __restrict__ un,
                                        int64_t c, int64_t si, int64_t sj, double thr,
                                        double* x, double* y) {
  const double u00 = u[c], u01 = u[c + sj], u10 = u[c + si], u11 = u[c + si + sj];
  const double n00 = un[c], n01 = un[c + sj], n10 = un[c + si], n11 = un[c + si + sj];
  const double AC = u00 - u01 + u11 - u10;
  const double GC = u01 - u00;
  const double BC = u10 - u00;
  const double DC = u00 - thr;
  const double AD = n00 - n01 + n11 - n10;
  const double GD = n01 - n00;
  const double BD = n10 - n00;
  const double DD = n00 - thr;
  const double Q = BC * AD - BD * AC;
  const double R = GC * AD - GD * AC;
  const double S = DC * AD - DD * AC;
  const double QOnR = Q / R;
  const double SOnR = S / R;
  const double T = AC * QOnR;
  const double Uq = AC * SOnR - BC + GC * QOnR;
  const double V = GC * SOnR - DC;
  const double disc = Uq * Uq - 4. * T * V;
  if (disc < 0) return false;  // a NaN goes on and fails the range checks below
  const double T2 = 2. * T;
  if (T2 == 0.) return false;
  const double sq = sqrt(disc);
  const double xn = (-Uq - sq) / T2;
  const double xp = (-Uq + sq) / T2;
  const double yn = -QOnR * xn - SOnR;
  const double yp = -QOnR * xp - SOnR;
  if (0 <= xn && xn <= 1 && 0 <= yn && yn <= 1) {
    *x = xn;
    *y = yn;
    return true;
  }
  if (0 <= xp && xp <= 1 && 0 <= yp && yp <= 1) {
    *x = xp;
    *y = yp;
    return true;
  }
  return false;
}

// _track_tip_line (spiral_wave_core_2d_tracker.py:211-249) over every slice
// k of the grid (one slice in 2D); each tip is appended as (i, j, k, x, y)
// through an atomic counter; the host restores the reference's (k, i, j)
// order.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_tips(const U* __restrict__ u_prev,
                                                 const U* __restrict__ u, int64_t ni,
                                                 int64_t nj, int64_t nk, double thr,
                                                 double* __restrict__ tips, int32_t cap,
                                                 int32_t* count) {
  const int64_t mi = ni - 10, mj = nj - 10;
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (mi <= 0 || mj <= 0 || q >= mi * mj * nk) return;
  const int64_t k = q % nk, r = q / nk, j = 5 + r % mj, i = 5 + r / mj;
  const int64_t sj = nk, si = nj * nk;
  const int64_t c = i * si + j * sj + k;
  int cnt = tip_cross(u_prev, c, si, sj, thr);
  if (cnt == 1) cnt += tip_cross(u, c, si, sj, thr);
  if (cnt != 2) return;
  double x, y;
  if (!tip_pos(u_prev, u, c, si, sj, thr, &x, &y)) return;
  const int32_t slot = atomicAdd(count, 1);
  if (slot >= cap) return;
  double* o = tips + (int64_t)slot * 5;
  o[0] = (double)i;
  o[1] = (double)j;
  o[2] = (double)k;
  o[3] = (double)j + y;  // the reference appends [j + y, i + x]
  o[4] = (double)i + x;
}

static unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int launch_lat(int u_f64, const void* u, int32_t* state, int32_t* steps, int32_t cap,
               double thr, int32_t step, int64_t n, int32_t* overflow, hipStream_t s) {
  if (n == 0) return FW_OK;
  if (u_f64)
    hipLaunchKernelGGL(k_lat<double>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                       static_cast<const double*>(u), state, steps, cap, thr, step, n, overflow);
  else
    hipLaunchKernelGGL(k_lat<float>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                       static_cast<const float*>(u), state, steps, cap, (float)thr, step, n,
                       overflow);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("lat launch failed: %s", hipGetErrorString(e));
    return FW_E_HIP;
  }
  return FW_OK;
}

int launch_tips(int u_f64, const void* u_prev, const void* u, const int64_t* shape, double thr,
                double* tips, int32_t cap, int32_t* count, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, count, (int64_t)1);
  const int64_t mi = shape[0] - 10, mj = shape[1] - 10;
  const int64_t n = (mi > 0 && mj > 0) ? mi * mj * shape[2] : 0;
  if (n > 0) {
    if (u_f64)
      hipLaunchKernelGGL(k_tips<double>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                         static_cast<const double*>(u_prev), static_cast<const double*>(u),
                         shape[0], shape[1], shape[2], thr, tips, cap, count);
    else
      hipLaunchKernelGGL(k_tips<float>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                         static_cast<const float*>(u_prev), static_cast<const float*>(u),
                         shape[0], shape[1], shape[2], thr, tips, cap, count);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("tips launch failed: %s", hipGetErrorString(e));
    return FW_E_HIP;
  }
  return FW_OK;
}

int launch_zero_f64(double* p, int64_t n, hipStream_t s) {
  if (n <= 0) return FW_OK;
  hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
  return hipGetLastError() == hipSuccess ? FW_OK : FW_E_HIP;
}

}  // namespace fw
