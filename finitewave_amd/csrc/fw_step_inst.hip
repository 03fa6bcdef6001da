// fw_step_inst.hip -- instantiates the fused step kernel for one
// (precision, model) pair, chosen at compile time by FW_PREC and FW_MODEL.
// The build compiles this file once per pair (FP64 builds with
// -ffp-contract=off, FP32 builds with -ffp-contract=on).
#include <type_traits>

#include "fw_internal.hpp"
#include "fw_kernels.hpp"
#include "fw_march3d.hpp"

#ifndef FW_PREC
#error "FW_PREC must be defined"
#endif
#ifndef FW_MODEL
#error "FW_MODEL must be defined"
#endif

#define FW_CAT4_(a, b, c, d) a##b##c##d
#define FW_CAT4(a, b, c, d) FW_CAT4_(a, b, c, d)
#define FW_LAUNCHER FW_CAT4(fw_step_launch_p, FW_PREC, _m, FW_MODEL)

namespace fw {

using Real = std::conditional_t<FW_PREC == 0, double, float>;
using UReal = UType<FW_MODEL, Real>;  // storage type of u / u_new

template <int NDIM, int ST>
static void launch(const fw_plan* plan, const KArgs<Real, UReal>& a, hipStream_t s) {
#if FW_PREC == 1 && FW_MODEL < 5  // the float-potential models
  if constexpr (NDIM == 3 && ST == FW_STENCIL_ANISO && std::is_same_v<UReal, float>) {
    // the FP32 19-point step that computes each face and square once
    // (fw_march3d.hpp), for fields uniform or tabulated along k
    MarchGeom g;
    if ((a.field_layout == kFieldsUniform || a.field_layout == kFieldsK) &&
        march_geom(a.ni, a.nj, a.nk, g)) {
      const dim3 grid(g.per_xcd * kXcds);
      if (a.field_layout == kFieldsUniform)
        hipLaunchKernelGGL((k_step3_march<FW_MODEL, kFieldsUniform>), grid, dim3(kBlock), 0, s,
                           a, g);
      else
        hipLaunchKernelGGL((k_step3_march<FW_MODEL, kFieldsK>), grid, dim3(kBlock), 0, s, a, g);
      return;
    }
  }
#endif
  if constexpr (ST != FW_STENCIL_WEIGHTS) {
    if (a.field_layout == kFieldsUniform) {
      hipLaunchKernelGGL((k_step<NDIM, ST, FW_MODEL, Real, FW_PREC, kFieldsUniform>),
                         dim3(plan->grid), dim3(kBlock), 0, s, a);
      return;
    }
    if constexpr (NDIM == 3) {
      if (a.field_layout == kFieldsK) {
        hipLaunchKernelGGL((k_step<NDIM, ST, FW_MODEL, Real, FW_PREC, kFieldsK>),
                           dim3(plan->grid), dim3(kBlock), 0, s, a);
        return;
      }
    }
  }
  hipLaunchKernelGGL((k_step<NDIM, ST, FW_MODEL, Real, FW_PREC, kFieldsGeneral>),
                     dim3(plan->grid), dim3(kBlock), 0, s, a);
}

static void fill_args(const fw_plan* plan, const void* u, void* u_new, const StatePtrs& st,
                      const TrackArgs& tr, KArgs<Real, UReal>& a) {
  const fw_plan_desc& d = plan->desc;
  a.u = static_cast<const UReal*>(u);
  a.u_new = static_cast<UReal*>(u_new);
  a.st = st;
  a.code = d.code;
  a.ni = d.shape[0];
  a.nj = d.shape[1];
  a.nk = d.ndim == 3 ? d.shape[2] : 1;
  a.N = plan->N;
  a.idx32 = a.N <= 0xffffffffll;
  a.div_jk = make_udiv(a.idx32 ? (uint32_t)(a.nj * a.nk) : 1u);
  a.div_k = make_udiv(a.idx32 ? (uint32_t)a.nk : 1u);
  a.nblocks = plan->nblocks;
  a.per_xcd = plan->per_xcd;
  for (int q = 0; q < 9; ++q) {
    if (q < d.n_fields) {
      a.f[q].mode = d.fields[q].mode;
      a.f[q].axis = d.fields[q].axis;
      a.f[q].value = d.fields[q].value;
      a.f[q].data = d.fields[q].data;
      a.f[q].fvalue = (float)d.fields[q].value;
      a.f[q].fdata = plan->ftab[q];
    } else {
      a.f[q] = Field{0, 0, 0.0, nullptr, 0.0f, nullptr};
    }
  }
  a.field_layout = plan->field_layout;
  if (FW_PREC == 0 && plan->field_layout == kFieldsUniform)
    for (int q = 0; q < d.n_fields; ++q) minor_quotients(d.fields[q].value, a.mq[q]);
  if (plan->field_layout == kFieldsK) {
    // every field as its table along the last axis, which the general
    // dispatch (the ECG kernel) reads the same way
    for (int q = 0; q < d.n_fields; ++q) {
      a.f[q].mode = 1;
      a.f[q].axis = 2;
      a.f[q].data = plan->ktab[q];
      a.f[q].fdata = plan->fktab[q];
    }
  }
  if (d.stencil == FW_STENCIL_WEIGHTS) a.f[0].mode = plan->K;
  a.w = d.weights;
  a.D = d.D_model;
  a.dtd = d.dt;
  a.dr2 = d.dr2;
  {
    int e2;
    const double mant = frexp(d.dr2, &e2);
    a.inv_dr2 = (mant == 0.5 && e2 > -1000 && e2 < 1000) ? ldexp(1.0, 1 - e2) : 0.0;
  }
  if (FW_PREC == 0 && plan->field_layout == kFieldsUniform && d.ndim == 2) {
    // the exact weights of a node whose 3x3 neighbourhood all conducts, by
    // the kernel's own functions (this FP64 unit is built without
    // contraction); the kernel takes them for every such node
    Nbr M;
    for (int q = 0; q < 27; ++q) M.m[q] = (q % 3) == 1;  // the dk = 0 plane
    const double f0 = d.fields[0].value, f1 = d.fields[1].value;
    double* W = a.wfull;
    if (d.stencil == FW_STENCIL_ANISO) {
      const double f2 = d.fields[2].value, f3 = d.fields[3].value;
      wts_aniso2d(W, M, f0, f0, f1, f1, f2, f2, f3, f3);
      scale_exact<9>(W, 4, a.D, a.dtd, a.dr2, a.inv_dr2);
    } else {
      W[0] = iso_comp(f0, 1, 1);
      W[1] = iso_comp(f1, 1, 1);
      W[3] = iso_comp(f1, 1, 1);
      W[4] = iso_comp(f0, 1, 1);
      W[2] = -(W[0] + W[1] + W[3] + W[4]);
      scale_exact<5>(W, 2, a.D, a.dtd, a.dr2, a.inv_dr2);
    }
  }
  a.dt = (Real)d.dt;
  a.s = (UReal)(d.D_model * d.dt / d.dr2);
  for (int q = 0; q < 64; ++q) a.P.v[q] = q < d.n_params ? d.params[q] : 0.0;
  complete_params(d.model, a.P);
  a.act = tr.act;
  a.act_thr = (UReal)tr.act_thr;
  a.act_on = tr.act_on && tr.act != nullptr;
  a.step = tr.step;
  a.step_in = tr.step_in;
  a.step_out = tr.step_out;
  a.n_probe = tr.n_probe;
  a.probe_idx = tr.probe_idx;
  a.probe_var = tr.probe_var;
  a.probe_buf = tr.probe_buf;
  a.probe_row = tr.probe_buf ? tr.probe_row : -1;
  a.n_stim = 0;
  a.stim_end = tr.stim_end;
  a.stim_end_in = tr.stim_end_in;
}

// the plan's stimulus schedule into the step kernel's arguments
static void fill_stims(const fw_plan* plan, KArgs<Real, UReal>& a) {
  a.n_stim = plan->n_stims;
  for (int x = 0; x < 3; ++x) {
    a.stim_lo[x] = INT32_MAX;
    a.stim_hi[x] = INT32_MIN;
  }
  for (int q = 0; q < plan->n_stims; ++q) {
    const fw_stim_sched& ss = plan->stims[q];
    StimRec& r = a.stim[q];
    for (int x = 0; x < 3; ++x) {
      r.lo[x] = (int32_t)ss.stim.lo[x];
      r.hi[x] = (int32_t)ss.stim.hi[x];
      a.stim_lo[x] = r.lo[x] < a.stim_lo[x] ? r.lo[x] : a.stim_lo[x];
      a.stim_hi[x] = r.hi[x] > a.stim_hi[x] ? r.hi[x] : a.stim_hi[x];
    }
    r.kind = ss.stim.kind;
    r.clamp = ss.stim.clamp;
    r.s_lo = (int32_t)ss.step_lo;
    r.s_hi = (int32_t)ss.step_hi;
    r.mask = ss.stim.mask;
    r.vals = ss.values;
    r.value = ss.stim.value;
    r.u_max = ss.stim.u_max;
  }
}

int FW_LAUNCHER(const fw_plan* plan, const void* u, void* u_new,
                const StatePtrs& st, const TrackArgs& tr, hipStream_t s) {
  const fw_plan_desc& d = plan->desc;
  KArgs<Real, UReal> a;
  fill_args(plan, u, u_new, st, tr, a);
  fill_stims(plan, a);
  const int st_kind = d.stencil;
  if (d.ndim == 2) {
    if (st_kind == 0) launch<2, 0>(plan, a, s);
    else if (st_kind == 1) launch<2, 1>(plan, a, s);
    else launch<2, 2>(plan, a, s);
  } else {
    if (st_kind == 0) launch<3, 0>(plan, a, s);
    else if (st_kind == 1) launch<3, 1>(plan, a, s);
    else launch<3, 2>(plan, a, s);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("step launch failed: %s", hipGetErrorString(e));
    return FW_E_HIP;
  }
  return FW_OK;
}

#define FW_ECG_LAUNCHER FW_CAT4(fw_ecg_launch_p, FW_PREC, _m, FW_MODEL)

template <int NDIM, int ST>
static void launch_ecg_k(const fw_plan* plan, const KArgs<Real, UReal>& a, const double* coords,
                         int32_t n, double dr, double* out, hipStream_t s) {
  hipLaunchKernelGGL((k_ecg<NDIM, ST, FW_MODEL, Real, FW_PREC>), dim3(plan->grid), dim3(kBlock),
                     0, s, a, coords, n, dr, out);
}

int FW_ECG_LAUNCHER(const fw_plan* plan, const void* u, const double* coords, int32_t n_points,
                    double dr, double* out, hipStream_t s) {
  const fw_plan_desc& d = plan->desc;
  KArgs<Real, UReal> a;
  TrackArgs none{};
  StatePtrs st{};
  fill_args(plan, u, nullptr, st, none, a);
  int rc = launch_zero_f64(out, n_points, s);
  if (rc) return rc;
  for (int32_t e0 = 0; e0 < n_points; e0 += kEcgPoints) {
    const int32_t m = n_points - e0 < kEcgPoints ? n_points - e0 : kEcgPoints;
    const double* cp = coords + 3 * e0;
    double* op = out + e0;
    if (d.ndim == 2) {
      if (d.stencil == 0) launch_ecg_k<2, 0>(plan, a, cp, m, dr, op, s);
      else if (d.stencil == 1) launch_ecg_k<2, 1>(plan, a, cp, m, dr, op, s);
      else launch_ecg_k<2, 2>(plan, a, cp, m, dr, op, s);
    } else {
      if (d.stencil == 0) launch_ecg_k<3, 0>(plan, a, cp, m, dr, op, s);
      else if (d.stencil == 1) launch_ecg_k<3, 1>(plan, a, cp, m, dr, op, s);
      else launch_ecg_k<3, 2>(plan, a, cp, m, dr, op, s);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("ecg launch failed: %s", hipGetErrorString(e));
    return FW_E_HIP;
  }
  return FW_OK;
}

}  // namespace fw
