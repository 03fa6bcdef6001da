// fw_batch.hpp -- many 2D tissues stepped by one launch per time step.
//
// Each tissue keeps its own HBM buffers (u/u_new pair, state arrays, node
// codes, activation-time map, probe trace) and a descriptor in a device
// table.  A block table maps every 256-node block of the launch to one
// tissue, so tissues of uneven size share a launch without padding, and the
// host admits or retires tissues by rewriting the block table between
// launches.  Fields are uniform per tissue (scalar conductivity, one fibre
// direction), held as four half-step values in the descriptor.  Each tissue
// counts its own local step (global step - step0), which selects its
// u/u_new orientation, gates its activation-time recording and indexes its
// probe trace.
#pragma once

#include "../../include/finitewave_amd.h"
#include "fw_kernels.hpp"

namespace fw {

// the public descriptor (include/finitewave_amd.h) is the device layout
typedef fw_tissue_desc TissueDesc;

template <typename R>
struct BatchArgs {
  const TissueDesc* __restrict__ T;
  const int32_t* __restrict__ btis;    // tissue of each logical block
  const int32_t* __restrict__ bstart;  // first local node of each logical block
  uint32_t nblocks, per_xcd;
  const int32_t* step_in;
  int32_t* step_out;
  int32_t step;
  double D, dtd, dr2, inv_dr2;
  R dt, s;
  Params P;
};

// u at a node, 2D, uniform fields: FP64 exact weights or FP32 flux form
template <typename R>
__device__ __forceinline__ R batch_diffuse(const TissueDesc& d, const R* __restrict__ u,
                                           const uint8_t* __restrict__ code, int64_t c,
                                           R uc, double D, double dtd, double dr2,
                                           double inv_dr2, R s) {
  const int64_t sN = d.nj;
  if (d.stencil == 0) {
    const int mN = code[c - sN] & 1, mS = code[c + sN] & 1;
    const int mW = code[c - 1] & 1, mE = code[c + 1] & 1;
    if constexpr (sizeof(R) == 8) {
      double W[5];
      W[0] = iso_comp(d.d[0], mN, mS);
      W[1] = iso_comp(d.d[1], mW, mE);
      W[3] = iso_comp(d.d[1], mE, mW);
      W[4] = iso_comp(d.d[0], mS, mN);
      W[2] = -(W[0] + W[1] + W[3] + W[4]);
      scale_exact<5>(W, 2, D, dtd, dr2, inv_dr2);
      return u[c - sN] * W[0] + u[c - 1] * W[1] + uc * W[2] + u[c + 1] * W[3] +
             u[c + sN] * W[4];
    } else {
      const R dx = (R)d.d[0], dy = (R)d.d[1];
      R fl = R(0);
      if (mN) fl += dx * (mS ? R(1) : R(2)) * (u[c - sN] - uc);
      if (mW) fl += dy * (mE ? R(1) : R(2)) * (u[c - 1] - uc);
      if (mE) fl += dy * (mW ? R(1) : R(2)) * (u[c + 1] - uc);
      if (mS) fl += dx * (mN ? R(1) : R(2)) * (u[c + sN] - uc);
      return uc + s * fl;
    }
  }
  Nbr M;
#pragma unroll
  for (int q = 0; q < 27; ++q) M.m[q] = 0;
#pragma unroll
  for (int di = -1; di <= 1; ++di)
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj) M.m[(di + 1) * 9 + (dj + 1) * 3 + 1] = code[c + di * sN + dj] & 1;
  const double dxx = d.d[0], dxy = d.d[1], dyx = d.d[2], dyy = d.d[3];
  if constexpr (sizeof(R) == 8) {
    double W[9];
    wts_aniso2d(W, M, dxx, dxx, dxy, dxy, dyx, dyx, dyy, dyy);
    scale_exact<9>(W, 4, D, dtd, dr2, inv_dr2);
    return u[c - sN - 1] * W[0] + u[c - sN] * W[1] + u[c - sN + 1] * W[2] + u[c - 1] * W[3] +
           uc * W[4] + u[c + 1] * W[5] + u[c + sN - 1] * W[6] + u[c + sN] * W[7] +
           u[c + sN + 1] * W[8];
  } else {
    const R uNW = u[c - sN - 1], uN = u[c - sN], uNE = u[c - sN + 1];
    const R uW = u[c - 1], uE = u[c + 1];
    const R uSW = u[c + sN - 1], uS = u[c + sN], uSE = u[c + sN + 1];
    const int mNW = M(-1, -1, 0), mN = M(-1, 0, 0), mNE = M(-1, 1, 0);
    const int mW = M(0, -1, 0), mE = M(0, 1, 0);
    const int mSW = M(1, -1, 0), mS = M(1, 0, 0), mSE = M(1, 1, 0);
    R fl = R(0);
    if (mN) fl += (R)dxx * (uN - uc);
    fl -= minor_flux<R>((R)dxy, mNW, uNW, mW, uW, mN, uN, 1, uc, mNE, uNE, mE, uE);
    if (mW) fl += (R)dyy * (uW - uc);
    fl -= minor_flux<R>((R)dyx, mNW, uNW, mN, uN, mW, uW, 1, uc, mSW, uSW, mS, uS);
    if (mE) fl += (R)dyy * (uE - uc);
    fl += minor_flux<R>((R)dyx, mNE, uNE, mN, uN, mE, uE, 1, uc, mSE, uSE, mS, uS);
    if (mS) fl += (R)dxx * (uS - uc);
    fl += minor_flux<R>((R)dxy, mSW, uSW, mW, uW, mS, uS, 1, uc, mSE, uSE, mE, uE);
    return uc + s * fl;
  }
}

template <int MODEL, typename R>
__global__ __launch_bounds__(kBlock) void k_batch_step(const BatchArgs<R> a) {
  const int64_t lb = logical_block(a.per_xcd);
  if (lb >= a.nblocks) return;
  const int t = a.btis[lb];
  const TissueDesc& d = a.T[t];
  const int64_t c = (int64_t)a.bstart[lb] + threadIdx.x;
  const int32_t step = a.step_in ? *a.step_in : a.step;
  if (a.step_out && lb == 0 && threadIdx.x == 0) *a.step_out = step + 1;
  if (c >= d.n) return;
  const int32_t ls = step - d.step0;
  if (ls < 0) return;
  const R* __restrict__ u = static_cast<const R*>((ls & 1) ? d.un : d.u);
  R* __restrict__ un = static_cast<R*>((ls & 1) ? d.u : d.un);
  const uint8_t code = d.code[c];
  const R uc = u[c];
  if (d.act && d.act_start >= 0 && ls >= d.act_start) {
    if (d.act[c] < 0 && uc > (R)d.act_thr) d.act[c] = ls;
  }
  if (d.probe && c == d.probe_idx && ls < d.probe_len) static_cast<R*>(d.probe)[ls] = uc;
  if ((code & 3) != 1) return;
  const int64_t i = c / d.nj, j = c - i * d.nj;
  if (i <= 0 || i >= d.ni - 1 || j <= 0 || j >= d.nj - 1) return;
  StatePtrs S;
#pragma unroll
  for (int k = 0; k < 24; ++k) S.p[k] = d.state[k];
  R val = batch_diffuse<R>(d, u, d.code, c, uc, a.D, a.dtd, a.dr2, a.inv_dr2, a.s);
  val = ionic_update<MODEL, R>(uc, val, S, c, a.dt, a.P);
  un[c] = val;
}

}  // namespace fw
