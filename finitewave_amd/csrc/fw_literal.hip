// fw_literal.hip -- launchers of the one-sweep kernels that mirror the
// reference's numba entry points (weights, diffusion, ionic), plus the stim
// and activation-time kernels.  Built with -ffp-contract=off.
#include "fw_internal.hpp"
#include "fw_kernels.hpp"

namespace fw {

static inline unsigned blocks_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}

static int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s launch failed: %s", what, hipGetErrorString(e));
    return FW_E_HIP;
  }
  return FW_OK;
}

int launch_weights(int ndim, int kind, double* w, const int8_t* m,
                   const double* const* d_dev, const int64_t* shape,
                   hipStream_t s) {
  const int64_t ni = shape[0], nj = shape[1], nk = ndim == 3 ? shape[2] : 1;
  const int64_t N = ni * nj * nk;
  if (N == 0) return FW_OK;
  dim3 g(blocks_for(N)), b(kBlock);
  DPtrs dp;
  const int nf = ndim == 2 ? (kind == FW_STENCIL_ISO ? 2 : 4) : (kind == FW_STENCIL_ISO ? 3 : 9);
  for (int q = 0; q < 9; ++q) dp.p[q] = q < nf ? d_dev[q] : nullptr;
  if (ndim == 2 && kind == FW_STENCIL_ISO)
    hipLaunchKernelGGL((k_weights<2, 0>), g, b, 0, s, w, m, dp, ni, nj, nk);
  else if (ndim == 2 && kind == FW_STENCIL_ANISO)
    hipLaunchKernelGGL((k_weights<2, 1>), g, b, 0, s, w, m, dp, ni, nj, nk);
  else if (ndim == 2 && kind == FW_STENCIL_SYM)
    hipLaunchKernelGGL((k_weights<2, 3>), g, b, 0, s, w, m, dp, ni, nj, nk);
  else if (ndim == 3 && kind == FW_STENCIL_ISO)
    hipLaunchKernelGGL((k_weights<3, 0>), g, b, 0, s, w, m, dp, ni, nj, nk);
  else if (ndim == 3 && kind == FW_STENCIL_ANISO)
    hipLaunchKernelGGL((k_weights<3, 1>), g, b, 0, s, w, m, dp, ni, nj, nk);
  else {
    set_error("compute_weights: unsupported ndim=%d kind=%d", ndim, kind);
    return FW_E_UNSUPPORTED;
  }
  return check_launch("compute_weights");
}

int launch_diffusion(int ndim, int K, double* u_new, const double* u,
                     const double* w, const int64_t* idx, int64_t n,
                     const int64_t* shape, hipStream_t s) {
  if (n == 0) return FW_OK;
  const int64_t nj = shape[1], nk = ndim == 3 ? shape[2] : 1;
  dim3 g(blocks_for(n)), b(kBlock);
  if (ndim == 2 && K == 5)
    hipLaunchKernelGGL((k_diffusion<2, 5>), g, b, 0, s, u_new, u, w, idx, n, nj, nk);
  else if (ndim == 2 && K == 9)
    hipLaunchKernelGGL((k_diffusion<2, 9>), g, b, 0, s, u_new, u, w, idx, n, nj, nk);
  else if (ndim == 3 && K == 7)
    hipLaunchKernelGGL((k_diffusion<3, 7>), g, b, 0, s, u_new, u, w, idx, n, nj, nk);
  else if (ndim == 3 && K == 19)
    hipLaunchKernelGGL((k_diffusion<3, 19>), g, b, 0, s, u_new, u, w, idx, n, nj, nk);
  else {
    set_error("diffusion: unsupported ndim=%d K=%d", ndim, K);
    return FW_E_UNSUPPORTED;
  }
  return check_launch("diffusion");
}

template <int MODEL>
static void ionic_t(int precision, void* u_new, const void* u, const StatePtrs& st,
                    const int64_t* idx, int64_t n, double dt, const Params& P,
                    hipStream_t s) {
  dim3 g(blocks_for(n)), b(kBlock);
  if (precision == FW_PREC_FP64)
    hipLaunchKernelGGL((k_ionic<MODEL, double>), g, b, 0, s, static_cast<double*>(u_new),
                       static_cast<const double*>(u), st, idx, n, dt, P);
  else
    hipLaunchKernelGGL((k_ionic<MODEL, float>), g, b, 0, s,
                       static_cast<UType<MODEL, float>*>(u_new),
                       static_cast<const UType<MODEL, float>*>(u), st, idx, n, (float)dt, P);
}

int launch_ionic(int model, int precision, void* u_new, const void* u,
                 const StatePtrs& st, const int64_t* idx, int64_t n, double dt,
                 const Params& P, hipStream_t s) {
  if (n == 0) return FW_OK;
  switch (model) {
    case 0: ionic_t<0>(precision, u_new, u, st, idx, n, dt, P, s); break;
    case 1: ionic_t<1>(precision, u_new, u, st, idx, n, dt, P, s); break;
    case 2: ionic_t<2>(precision, u_new, u, st, idx, n, dt, P, s); break;
    case 3: ionic_t<3>(precision, u_new, u, st, idx, n, dt, P, s); break;
    case 4: ionic_t<4>(precision, u_new, u, st, idx, n, dt, P, s); break;
    case 5: ionic_t<5>(precision, u_new, u, st, idx, n, dt, P, s); break;
    case 6: ionic_t<6>(precision, u_new, u, st, idx, n, dt, P, s); break;
    case 7: ionic_t<7>(precision, u_new, u, st, idx, n, dt, P, s); break;
    default:
      set_error("ionic: unknown model %d", model);
      return FW_E_INVALID;
  }
  return check_launch("ionic");
}

int launch_stim(int u_f64, void* u, const uint8_t* code, const fw_stim_desc* st,
                const int64_t* shape, hipStream_t s, const double* value_dev) {
  int64_t lo[3], ext[3];
  for (int a = 0; a < 3; ++a) {
    int64_t l = st->lo[a] < 0 ? 0 : st->lo[a];
    int64_t h = st->hi[a] > shape[a] ? shape[a] : st->hi[a];
    lo[a] = l;
    ext[a] = h > l ? h - l : 0;
  }
  const int64_t n = ext[0] * ext[1] * ext[2];
  if (n == 0) return FW_OK;
  dim3 g(blocks_for(n)), b(kBlock);
  if (u_f64)
    hipLaunchKernelGGL((k_stim<double>), g, b, 0, s, static_cast<double*>(u), code, st->mask,
                       lo[0], lo[1], lo[2], ext[0], ext[1], ext[2], shape[1], shape[2],
                       st->kind, st->value, value_dev, st->clamp, st->u_max);
  else
    hipLaunchKernelGGL((k_stim<float>), g, b, 0, s, static_cast<float*>(u), code, st->mask,
                       lo[0], lo[1], lo[2], ext[0], ext[1], ext[2], shape[1], shape[2],
                       st->kind, st->value, value_dev, st->clamp, st->u_max);
  return check_launch("stim");
}

int launch_act(int u_f64, int32_t* act, const void* u, int64_t n, double thr,
               int32_t step, hipStream_t s) {
  if (n == 0) return FW_OK;
  dim3 g(blocks_for(n)), b(kBlock);
  if (u_f64)
    hipLaunchKernelGGL((k_act<double>), g, b, 0, s, act, static_cast<const double*>(u), n,
                       thr, step);
  else
    hipLaunchKernelGGL((k_act<float>), g, b, 0, s, act, static_cast<const float*>(u), n,
                       (float)thr, step);
  return check_launch("act_time");
}

}  // namespace fw
