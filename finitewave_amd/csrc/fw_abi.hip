// fw_abi.hip -- the extern "C" surface declared in include/finitewave_amd.h.
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "fw_batch.hpp"
#include "fw_internal.hpp"

namespace fw {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

#define FW_DECL_LAUNCHER(p, m)                                                 \
  int fw_step_launch_p##p##_m##m(const fw_plan*, const void*, void*,           \
                                 const StatePtrs&, const TrackArgs&, hipStream_t);
FW_DECL_LAUNCHER(0, 0) FW_DECL_LAUNCHER(0, 1) FW_DECL_LAUNCHER(0, 2) FW_DECL_LAUNCHER(0, 3)
FW_DECL_LAUNCHER(0, 4) FW_DECL_LAUNCHER(0, 5) FW_DECL_LAUNCHER(0, 6) FW_DECL_LAUNCHER(0, 7)
FW_DECL_LAUNCHER(1, 0) FW_DECL_LAUNCHER(1, 1) FW_DECL_LAUNCHER(1, 2) FW_DECL_LAUNCHER(1, 3)
FW_DECL_LAUNCHER(1, 4) FW_DECL_LAUNCHER(1, 5) FW_DECL_LAUNCHER(1, 6) FW_DECL_LAUNCHER(1, 7)

StepLauncher step_launcher(int precision, int model) {
  static const StepLauncher table[2][8] = {
      {fw_step_launch_p0_m0, fw_step_launch_p0_m1, fw_step_launch_p0_m2, fw_step_launch_p0_m3,
       fw_step_launch_p0_m4, fw_step_launch_p0_m5, fw_step_launch_p0_m6, fw_step_launch_p0_m7},
      {fw_step_launch_p1_m0, fw_step_launch_p1_m1, fw_step_launch_p1_m2, fw_step_launch_p1_m3,
       fw_step_launch_p1_m4, fw_step_launch_p1_m5, fw_step_launch_p1_m6, fw_step_launch_p1_m7}};
  if (precision < 0 || precision > 1 || model < 0 || model > 7) return nullptr;
  return table[precision][model];
}

#define FW_DECL_ECG(p, m)                                                       \
  int fw_ecg_launch_p##p##_m##m(const fw_plan*, const void*, const double*, int32_t, double, \
                                double*, hipStream_t);
FW_DECL_ECG(0, 0) FW_DECL_ECG(0, 1) FW_DECL_ECG(0, 2) FW_DECL_ECG(0, 3)
FW_DECL_ECG(0, 4) FW_DECL_ECG(0, 5) FW_DECL_ECG(0, 6) FW_DECL_ECG(0, 7)
FW_DECL_ECG(1, 0) FW_DECL_ECG(1, 1) FW_DECL_ECG(1, 2) FW_DECL_ECG(1, 3)
FW_DECL_ECG(1, 4) FW_DECL_ECG(1, 5) FW_DECL_ECG(1, 6) FW_DECL_ECG(1, 7)

EcgLauncher ecg_launcher(int precision, int model) {
  static const EcgLauncher table[2][8] = {
      {fw_ecg_launch_p0_m0, fw_ecg_launch_p0_m1, fw_ecg_launch_p0_m2, fw_ecg_launch_p0_m3,
       fw_ecg_launch_p0_m4, fw_ecg_launch_p0_m5, fw_ecg_launch_p0_m6, fw_ecg_launch_p0_m7},
      {fw_ecg_launch_p1_m0, fw_ecg_launch_p1_m1, fw_ecg_launch_p1_m2, fw_ecg_launch_p1_m3,
       fw_ecg_launch_p1_m4, fw_ecg_launch_p1_m5, fw_ecg_launch_p1_m6, fw_ecg_launch_p1_m7}};
  if (precision < 0 || precision > 1 || model < 0 || model > 7) return nullptr;
  return table[precision][model];
}

static void derive_tp06(Params& P) {
  // TP06 parameter order (tp06_2d.py:1073 call)
  const double ko = P.v[0], cao = P.v[1], nao = P.v[2], Vc = P.v[3], Vsr = P.v[4],
               Vss = P.v[5], R = P.v[24], F = P.v[25], T = P.v[26], KmK = P.v[34],
               KmNai = P.v[40], KmCa = P.v[41];
  double* d = P.v + kTp06Derived;
  d[kInvVcF] = 1.0 / (Vc * F);
  d[kInvVssF2] = 1.0 / (2 * Vss * F);
  d[kFFoverRT] = F * F / (R * T);
  d[kRT] = R * T;
  d[kInvRT] = 1.0 / (R * T);
  d[kInvNaCaK] = 1.0 / (KmNai * KmNai * KmNai + nao * nao * nao);
  d[kInvKmCaCao] = 1.0 / (KmCa + cao);
  d[kKoOverKoKmK] = ko / (ko + KmK);
  d[kSqrtKo54] = sqrt(ko / 5.4);
  d[kVcOverVss] = Vc / Vss;
  d[kVsrOverVss] = Vsr / Vss;
}

void complete_params(int model, Params& P) {
  if (model == 6) derive_tp06(P);
  if (model == 2) {
    // Mitchell-Schaeffer, float path: 1 / tau_open, 1 / tau_close
    P.v[5] = 1.0 / P.v[1];
    P.v[6] = 1.0 / P.v[0];
  }
  for (int q = 0; q < 64; ++q) P.f[q] = (float)P.v[q];  // round to nearest
}

__global__ void k_set_word(int32_t* p, int32_t v) { *p = v; }

// graph replay words: the step the next launch reads, and the end of the
// steps whose scheduled stimuli the kernels apply
__global__ void k_set_run_words(int32_t* p, int32_t step0, int32_t stim_end) {
  p[0] = step0;
  p[2] = stim_end;
}

__global__ void k_to_float(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = (float)src[q];
}

// The quotients minor_exact forms from a field value d and the counts
// mh, ml in {3, 4}, in its operation order: -d / ml, d / mh and
// d * (1 / mh - 1 / ml).  This unit is built without contraction.
void minor_quotients(double d, MinorQ& T) {
  for (int h = 0; h < 2; ++h) {
    const double dm = h ? 4.0 : 3.0;
    T.nq[h] = (-d) / dm;
    T.pq[h] = d / dm;
    for (int l = 0; l < 2; ++l) {
      const double dl = l ? 4.0 : 3.0;
      const double c = 1.0 / dm - 1.0 / dl;
      T.dc[h][l] = d * c;
    }
  }
}

// A field as a table along the last axis (n entries): a copy of its table,
// or its uniform value repeated (src null); in double and in float.
__global__ void k_fill_ktab(const double* __restrict__ src, double value, double* __restrict__ kd,
                            float* __restrict__ kf, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double v = src ? src[q] : value;
  kd[q] = v;
  kf[q] = (float)v;
}

// Admission check of a batched tissue's codes: bit 0 of *bad when an updated
// node (code & 3 == 1) lies on the outer layer, bit 1 when a node is frozen
// (code bit 1).  Grid-stride over all n = ni * nj nodes.
__global__ void k_check_codes(const uint8_t* __restrict__ code, int64_t ni, int64_t nj,
                              int32_t* bad) {
  const int64_t n = ni * nj;
  int found = 0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t cd = code[c];
    if (cd & 2) found |= 2;
    if ((cd & 3) == 1) {
      const int64_t i = c / nj, j = c - i * nj;
      if (i == 0 || i == ni - 1 || j == 0 || j == nj - 1) found |= 1;
    }
  }
  if (found) atomicOr(bad, found);
}

// flags[t] |= 1 for every tissue t of the block table whose potential
// before global step `step` holds a NaN or an infinity: logical block lb
// covers `per` nodes of tissue btis[lb] from node bstart[lb].
template <typename U>
__global__ void k_batch_nonfinite(const fw::BatchTissue* __restrict__ T,
                                  const int32_t* __restrict__ btis,
                                  const int32_t* __restrict__ bstart, int32_t per,
                                  int32_t step, int32_t* flags) {
  const int32_t t = btis[blockIdx.x], start = bstart[blockIdx.x];
  const fw::BatchTissue& d = T[t];
  const int32_t ls = step - d.step0;
  if (ls < 0) return;
  const U* __restrict__ u = (const U*)((ls & 1) ? d.un : d.u);
  const int32_t end = start + per < d.n ? start + per : d.n;
  bool bad = false;
  for (int32_t c = start + (int32_t)threadIdx.x; c < end; c += (int32_t)blockDim.x)
    bad |= !__builtin_isfinite(u[c]);
  if (bad) atomicOr(flags + t, 1);
}

__global__ void k_fill_i32(int32_t* p, int32_t n, int32_t v) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = v;
}

// the seed mixed once on the host (SplitMix64's output function, below)
static uint64_t host_splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// SplitMix64's output function (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): a bijective 64-bit mix, here
// applied to a counter (node or block index) and a mixed seed.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Mesh codes of the region [lo, hi) of a C-order grid (shape n[3], 2D:
// n[2] = 1): 2 where the node's (or its block's) draw u in [0, 1) is
// <= density, else 1; grid-stride over the region's nodes.
__global__ void k_fibrosis(int8_t* __restrict__ mesh, int64_t n1, int64_t n2, int64_t lo0,
                           int64_t lo1, int64_t lo2, int64_t r1, int64_t r2, int64_t count,
                           int64_t b0, int64_t b1, int64_t b2, int64_t nb1, int64_t nb2,
                           int32_t diffuse, double density, uint64_t seed_mix) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < count;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ri = r / (r1 * r2), rj = (r / r2) % r1, rk = r % r2;
    const int64_t c = ((lo0 + ri) * n1 + (lo1 + rj)) * n2 + (lo2 + rk);
    const uint64_t key =
        diffuse ? (uint64_t)c : (uint64_t)(((ri / b0) * nb1 + rj / b1) * nb2 + rk / b2);
    const double u = (double)(splitmix64(key ^ seed_mix) >> 11) * 0x1.0p-53;
    mesh[c] = u <= density ? 2 : 1;
  }
}

}  // namespace fw

using namespace fw;

static hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

static StatePtrs make_state(void* const* state, int n) {
  StatePtrs st;
  for (int k = 0; k < FW_MAX_STATE; ++k) st.p[k] = k < n ? state[k] : nullptr;
  return st;
}

extern "C" {

int fw_abi_version(void) { return FW_ABI_VERSION; }

const char* fw_last_error(void) { return g_err; }

int fw_model_info(int32_t model, int32_t* n_state, int32_t* n_params) {
  if (model < 0 || model >= FW_N_MODELS) {
    set_error("unknown model %d", model);
    return FW_E_INVALID;
  }
  if (n_state) *n_state = model_n_state(model);
  if (n_params) *n_params = model_n_params(model);
  return FW_OK;
}

int fw_u_is_f64(int32_t model, int32_t precision) {
  if (model < 0 || model >= FW_N_MODELS ||
      (precision != FW_PREC_FP64 && precision != FW_PREC_FP32)) {
    set_error("fw_u_is_f64: bad arguments");
    return FW_E_INVALID;
  }
  return u_is_f64(model, precision) ? 1 : 0;
}

int fw_state_is_f64(int32_t model, int32_t precision, int32_t k) {
  if (model < 0 || model >= FW_N_MODELS || k < 0 || k >= model_n_state(model)) return FW_E_INVALID;
  return state_is_f64(model, precision, k) ? 1 : 0;
}

int fw_state_offset(int32_t model, int32_t precision, int32_t k, double* offset) {
  if (model < 0 || model >= FW_N_MODELS || k < 0 || k >= model_n_state(model) ||
      (precision != FW_PREC_FP64 && precision != FW_PREC_FP32) || !offset) {
    set_error("fw_state_offset: bad arguments");
    return FW_E_INVALID;
  }
  *offset = state_offset(model, precision, k);
  return FW_OK;
}

int fw_compute_weights(int32_t ndim, int32_t kind, double* w, const int8_t* m,
                       const double* const* d, const int64_t* shape, void* stream) {
  if (!w || !m || !d || !shape || (ndim != 2 && ndim != 3)) {
    set_error("fw_compute_weights: bad arguments");
    return FW_E_INVALID;
  }
  return launch_weights(ndim, kind, w, m, d, shape, as_stream(stream));
}

int fw_diffusion(int32_t ndim, int32_t K, double* u_new, const double* u, const double* w,
                 const int64_t* indexes, int64_t n_indexes, const int64_t* shape,
                 void* stream) {
  if (!u_new || !u || !w || (!indexes && n_indexes) || !shape || n_indexes < 0) {
    set_error("fw_diffusion: bad arguments");
    return FW_E_INVALID;
  }
  return launch_diffusion(ndim, K, u_new, u, w, indexes, n_indexes, shape, as_stream(stream));
}

int fw_ionic(int32_t model, int32_t precision, void* u_new, const void* u, void* const* state,
             int32_t n_state, const int64_t* indexes, int64_t n_indexes, double dt,
             const double* params, int32_t n_params, void* stream) {
  if (model < 0 || model >= FW_N_MODELS) {
    set_error("fw_ionic: unknown model %d", model);
    return FW_E_INVALID;
  }
  if (n_state != model_n_state(model) || n_params != model_n_params(model)) {
    set_error("fw_ionic: model %d takes %d state arrays and %d parameters (got %d, %d)", model,
              model_n_state(model), model_n_params(model), n_state, n_params);
    return FW_E_INVALID;
  }
  if (precision != FW_PREC_FP64 && precision != FW_PREC_FP32) {
    set_error("fw_ionic: bad precision %d", precision);
    return FW_E_INVALID;
  }
  if (!u_new || !u || !state || (!indexes && n_indexes) || !params || n_indexes < 0) {
    set_error("fw_ionic: null argument");
    return FW_E_INVALID;
  }
  for (int k = 0; k < n_state; ++k)
    if (!state[k]) {
      set_error("fw_ionic: state %d is null", k);
      return FW_E_INVALID;
    }
  Params P;
  for (int q = 0; q < FW_MAX_PARAMS; ++q) P.v[q] = q < n_params ? params[q] : 0.0;
  complete_params(model, P);
  return launch_ionic(model, precision, u_new, u, make_state(state, n_state), indexes,
                      n_indexes, dt, P, as_stream(stream));
}

// The half-step fields a stencil reads: ISO ndim, ANISO 4 (2D) or 9 (3D).
static int needed_fields(const fw_plan_desc& d) {
  return d.stencil == FW_STENCIL_WEIGHTS ? 0 : d.stencil == FW_STENCIL_ISO ? d.ndim
                                                : (d.ndim == 2 ? 4 : 9);
}

// Validates a plan description and fills the host part of a plan (no device
// memory): the description, node count, launch geometry and field layout.
// `who` names the entry point in error messages.
static int plan_host_init(const fw_plan_desc* desc, fw_plan* p, const char* who) {
  const fw_plan_desc& d = *desc;
  if (d.ndim != 2 && d.ndim != 3) {
    set_error("%s: ndim must be 2 or 3 (got %d)", who, d.ndim);
    return FW_E_INVALID;
  }
  if (d.model < 0 || d.model >= FW_N_MODELS) {
    set_error("%s: unknown model %d", who, d.model);
    return FW_E_INVALID;
  }
  if (d.precision != FW_PREC_FP64 && d.precision != FW_PREC_FP32) {
    set_error("%s: bad precision %d", who, d.precision);
    return FW_E_INVALID;
  }
  if (d.stencil < FW_STENCIL_ISO || d.stencil > FW_STENCIL_WEIGHTS) {
    set_error("%s: bad stencil %d", who, d.stencil);
    return FW_E_INVALID;
  }
  const int need_fields = needed_fields(d);
  if (d.n_fields < need_fields || d.n_fields > FW_MAX_FIELDS) {
    set_error("%s: stencil needs %d half-step fields (got %d)", who, need_fields, d.n_fields);
    return FW_E_INVALID;
  }
  for (int q = 0; q < need_fields; ++q) {
    const fw_field& f = d.fields[q];
    if (f.mode < FW_FIELD_UNIFORM || f.mode > FW_FIELD_FULL ||
        (f.mode != FW_FIELD_UNIFORM && !f.data) ||
        (f.mode == FW_FIELD_AXIS && (f.axis < 0 || f.axis >= d.ndim))) {
      set_error("%s: field %d is malformed", who, q);
      return FW_E_INVALID;
    }
  }
  if (d.stencil == FW_STENCIL_WEIGHTS && !d.weights) {
    set_error("%s: FW_STENCIL_WEIGHTS needs a weights array", who);
    return FW_E_INVALID;
  }
  if (d.n_params != model_n_params(d.model)) {
    set_error("%s: model %d takes %d parameters (got %d)", who, d.model,
              model_n_params(d.model), d.n_params);
    return FW_E_INVALID;
  }
  if (!d.code) {
    set_error("%s: code array is null", who);
    return FW_E_INVALID;
  }
  if (!(d.dt > 0) || !(d.dr2 > 0)) {
    set_error("%s: dt and dr2 must be positive", who);
    return FW_E_INVALID;
  }
  int64_t N = 1;
  for (int a = 0; a < d.ndim; ++a) {
    if (d.shape[a] < 3) {
      set_error("%s: every axis needs at least 3 nodes (axis %d has %lld)", who, a,
                (long long)d.shape[a]);
      return FW_E_INVALID;
    }
    N *= d.shape[a];
  }
  if (N > ((int64_t)1 << 40)) {
    set_error("%s: grid too large", who);
    return FW_E_INVALID;
  }
  p->desc = d;
  if (d.ndim == 2) p->desc.shape[2] = 1;
  p->N = N;
  p->nblocks = (uint32_t)((N + kBlock - 1) / kBlock);
  p->per_xcd = (p->nblocks + kXcds - 1) / kXcds;
  p->grid = p->per_xcd * kXcds;
  p->K = d.ndim == 2 ? 9 : 19;
  // the field layout: every field uniform (kFieldsUniform), or in 3D every
  // field uniform or tabulated along the last axis (kFieldsK, layered
  // fibres); such plans get their step kernel without the per-field mode
  // dispatch
  p->field_layout = fw::kFieldsGeneral;
  if (d.stencil != FW_STENCIL_WEIGHTS) {
    bool uni = true, along_k = d.ndim == 3;
    for (int q = 0; q < need_fields; ++q) {
      const fw_field& f = d.fields[q];
      if (f.mode != FW_FIELD_UNIFORM) uni = false;
      if (!(f.mode == FW_FIELD_UNIFORM || (f.mode == FW_FIELD_AXIS && f.axis == 2)))
        along_k = false;
    }
    if (uni) p->field_layout = fw::kFieldsUniform;
    else if (along_k) p->field_layout = fw::kFieldsK;
  }
  return FW_OK;
}

// Whether a plan's step kernel reads float copies of its tabulated fields
// (FP32 plans whose potential is float)
static bool wants_float_tables(const fw_plan_desc& d) {
  return d.precision == FW_PREC_FP32 && !u_is_f64(d.model, d.precision);
}

int fw_plan_create(const fw_plan_desc* desc, fw_plan** out) {
  if (!desc || !out) {
    set_error("fw_plan_create: null argument");
    return FW_E_INVALID;
  }
  *out = nullptr;
  const fw_plan_desc& d = *desc;
  fw_plan* p = new (std::nothrow) fw_plan();
  if (!p) {
    set_error("fw_plan_create: out of host memory");
    return FW_E_NOMEM;
  }
  int rc = plan_host_init(desc, p, "fw_plan_create");
  if (rc) {
    delete p;
    return rc;
  }
  const int need_fields = needed_fields(d);
  const int64_t N = p->N;
  hipError_t e = hipMalloc(&p->step_words, 4 * sizeof(int32_t));
  if (e != hipSuccess) {
    set_error("fw_plan_create: hipMalloc failed: %s", hipGetErrorString(e));
    delete p;
    return FW_E_NOMEM;
  }
  // a kernel computing the diffusion in float reads float copies of the
  // tables: 4-byte loads and no conversion per node
  if (wants_float_tables(d)) {
    bool synced = false;
    for (int q = 0; q < need_fields; ++q) {
      const fw_field& f = d.fields[q];
      if (f.mode == FW_FIELD_UNIFORM) continue;
      const int64_t len = f.mode == FW_FIELD_AXIS ? d.shape[f.axis] : N;
      if (!synced) {
        (void)hipDeviceSynchronize();  // the tables' uploads on any stream
        synced = true;
      }
      e = hipMalloc(&p->ftab[q], len * sizeof(float));
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_to_float, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, 0,
                           f.data, p->ftab[q], len);
      if (e == hipSuccess) e = hipGetLastError();
      if (e != hipSuccess) {
        set_error("fw_plan_create: float field table: %s", hipGetErrorString(e));
        fw_plan_destroy(p);
        return FW_E_NOMEM;
      }
    }
    if (synced) {
      e = hipDeviceSynchronize();
      if (e != hipSuccess) {
        set_error("fw_plan_create: float field tables: %s", hipGetErrorString(e));
        fw_plan_destroy(p);
        return FW_E_HIP;
      }
    }
  }
  if (p->field_layout == fw::kFieldsK) {
    // each field as a table along k, uniform ones filled with their value
    const int64_t nk = d.shape[2];
    (void)hipDeviceSynchronize();  // the tables' uploads on any stream
    for (int q = 0; q < need_fields && e == hipSuccess; ++q) {
      const fw_field& f = d.fields[q];
      e = hipMalloc(&p->ktab[q], nk * sizeof(double));
      if (e == hipSuccess) e = hipMalloc(&p->fktab[q], nk * sizeof(float));
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_fill_ktab, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, 0,
                           f.mode == FW_FIELD_AXIS ? f.data : nullptr, f.value, p->ktab[q],
                           p->fktab[q], nk);
      if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      set_error("fw_plan_create: field tables along k: %s", hipGetErrorString(e));
      fw_plan_destroy(p);
      return FW_E_NOMEM;
    }
  }
  *out = p;
  return FW_OK;
}

// the K of a FW_STENCIL_WEIGHTS plan: 5/9 (2D) or 7/19 (3D) weights per node
int fw_plan_set_weights_k(fw_plan* plan, int32_t K) {
  if (!plan) return FW_E_INVALID;
  const int nd = plan->desc.ndim;
  if (!((nd == 2 && (K == 5 || K == 9)) || (nd == 3 && (K == 7 || K == 19)))) {
    set_error("fw_plan_set_weights_k: K=%d does not fit ndim=%d", K, nd);
    return FW_E_INVALID;
  }
  plan->K = K;
  return FW_OK;
}

// Drops every captured step graph of a plan.
static void drop_graphs(fw_plan* plan) {
  for (auto& gs : plan->graphs) {
    for (auto& g : gs.g) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
    gs.u = gs.un = nullptr;
    gs.last_use = 0;
  }
  plan->graph_keyed = 0;
}

int fw_plan_destroy(fw_plan* plan) {
  if (!plan) return FW_OK;
  drop_graphs(plan);
  if (plan->step_words) (void)hipFree(plan->step_words);
  for (int q = 0; q < FW_MAX_FIELDS; ++q) {
    if (plan->ftab[q]) (void)hipFree(plan->ftab[q]);
    if (plan->ktab[q]) (void)hipFree(plan->ktab[q]);
    if (plan->fktab[q]) (void)hipFree(plan->fktab[q]);
  }
  delete plan;
  return FW_OK;
}

// a step number as the kernels hold it
static int32_t clamp_step(int64_t s) {
  return s < INT32_MIN ? INT32_MIN : (s > INT32_MAX ? INT32_MAX : (int32_t)s);
}

static int check_run_args(fw_plan* plan, void* u, void* u_new, void* const* state,
                          int32_t n_steps, const fw_track_desc* track) {
  if (!plan || !u || !u_new || !state || n_steps < 0) {
    set_error("fw_run: bad arguments");
    return FW_E_INVALID;
  }
  const int ns = model_n_state(plan->desc.model);
  for (int k = 0; k < ns; ++k)
    if (!state[k]) {
      set_error("fw_run: state %d is null", k);
      return FW_E_INVALID;
    }
  if (track && track->n_probe > 0 &&
      (!track->probe_idx || !track->probe_var || !track->probe_buf)) {
    set_error("fw_run: probes need idx, var and buf");
    return FW_E_INVALID;
  }
  return FW_OK;
}

// n_steps plain launches of the step kernel (arguments already checked)
static int run_plain(const fw_plan* plan, void* u, void* u_new, void* const* state,
                     int64_t step0, int32_t n_steps, const fw_track_desc* track, void* stream) {
  int rc = FW_OK;
  StepLauncher L = step_launcher(plan->desc.precision, plan->desc.model);
  const StatePtrs st = make_state(state, model_n_state(plan->desc.model));
  void* a = u;
  void* b = u_new;
  const int32_t stim_end = track ? clamp_step(track->stim_end) : 0;
  for (int32_t s = 0; s < n_steps; ++s) {
    TrackArgs tr{};
    tr.step = (int32_t)(step0 + s);
    tr.stim_end = stim_end;
    if (track) {
      tr.act = track->act;
      tr.act_thr = track->act_thr;
      tr.act_on = track->act && (!track->act_flags || track->act_flags[s]);
      tr.n_probe = track->n_probe;
      tr.probe_idx = track->probe_idx;
      tr.probe_var = track->probe_var;
      tr.probe_buf = track->probe_buf;
      tr.probe_row = track->probe_rows ? track->probe_rows[s] : -1;
    } else {
      tr.probe_row = -1;
    }
    rc = L(plan, a, b, st, tr, as_stream(stream));
    if (rc) return rc;
    void* t = a;
    a = b;
    b = t;
  }
  return FW_OK;
}

int fw_run(fw_plan* plan, void* u, void* u_new, void* const* state, int64_t step0,
           int32_t n_steps, const fw_track_desc* track, void* stream) {
  int rc = check_run_args(plan, u, u_new, state, n_steps, track);
  if (rc) return rc;
  return run_plain(plan, u, u_new, state, step0, n_steps, track, stream);
}

// ---------------------------------------------------------------------------
// Stepping without allocation (fw_run_noalloc)
// ---------------------------------------------------------------------------

// The staging memory of fw_run_noalloc: the float copies and along-k tables
// of a plan's half-step fields.  One copy per GPU, held by the loaded code
// object until the process ends.  In host code the name is a stand-in, not a
// device address: host code uses only the address hipGetSymbolAddress gives
// for the current device (scratch_for_current_device).
static __device__ __align__(256) unsigned char fw_scratch[FW_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices];

// The current device's address of fw_scratch, looked up on the first call on
// that device and kept
static int scratch_for_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("fw_run_noalloc: no usable current device (%s, %d)", hipGetErrorString(e), dev);
    return FW_E_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(fw_scratch));
    if (e != hipSuccess || !p) {
      set_error("fw_run_noalloc: hipGetSymbolAddress failed on device %d: %s", dev,
                hipGetErrorString(e));
      return FW_E_HIP;
    }
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return FW_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int fw_run_noalloc(const fw_plan_desc* desc, void* u, void* u_new, void* const* state,
                   int64_t step0, int32_t n_steps, const fw_track_desc* track, void* stream) {
  if (!desc) {
    set_error("fw_run_noalloc: null plan description");
    return FW_E_INVALID;
  }
  fw_plan p{};
  int rc = plan_host_init(desc, &p, "fw_run_noalloc");
  if (rc) return rc;
  rc = check_run_args(&p, u, u_new, state, n_steps, track);
  if (rc) return rc;
  const fw_plan_desc& d = p.desc;
  const int need = needed_fields(d);
  // offsets of the tables in the scratch array, then its size
  size_t off_f[FW_MAX_FIELDS] = {}, off_kd[FW_MAX_FIELDS] = {}, off_kf[FW_MAX_FIELDS] = {};
  int64_t len_f[FW_MAX_FIELDS] = {};
  size_t used = 0;
  const bool ftabs = wants_float_tables(d);
  for (int q = 0; q < need; ++q) {
    const fw_field& f = d.fields[q];
    if (!ftabs || f.mode == FW_FIELD_UNIFORM) continue;
    len_f[q] = f.mode == FW_FIELD_AXIS ? d.shape[f.axis] : p.N;
    off_f[q] = used;
    used = align256(used + (size_t)len_f[q] * sizeof(float));
  }
  const int64_t nk = d.shape[2];
  if (p.field_layout == kFieldsK) {
    for (int q = 0; q < need; ++q) {
      off_kd[q] = used;
      used = align256(used + (size_t)nk * sizeof(double));
      off_kf[q] = used;
      used = align256(used + (size_t)nk * sizeof(float));
    }
  }
  if (used > FW_NOALLOC_SCRATCH_BYTES) {
    set_error("fw_run_noalloc: the plan's field tables need %zu bytes, more than the %u bytes "
              "of static scratch; use fw_plan_create",
              used, (unsigned)FW_NOALLOC_SCRATCH_BYTES);
    return FW_E_INVALID;
  }
  hipStream_t s = as_stream(stream);
  if (used > 0) {
    void* base = nullptr;
    rc = scratch_for_current_device(&base);
    if (rc) return rc;
    unsigned char* b = static_cast<unsigned char*>(base);
    for (int q = 0; q < need; ++q) {
      if (len_f[q] > 0) {
        p.ftab[q] = reinterpret_cast<float*>(b + off_f[q]);
        hipLaunchKernelGGL(k_to_float, dim3((unsigned)((len_f[q] + 255) / 256)), dim3(256), 0,
                           s, d.fields[q].data, p.ftab[q], len_f[q]);
      }
      if (p.field_layout == kFieldsK) {
        const fw_field& f = d.fields[q];
        p.ktab[q] = reinterpret_cast<double*>(b + off_kd[q]);
        p.fktab[q] = reinterpret_cast<float*>(b + off_kf[q]);
        hipLaunchKernelGGL(k_fill_ktab, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s,
                           f.mode == FW_FIELD_AXIS ? f.data : nullptr, f.value, p.ktab[q],
                           p.fktab[q], nk);
      }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("fw_run_noalloc: table launch failed: %s", hipGetErrorString(e));
      return FW_E_HIP;
    }
  }
  return run_plain(&p, u, u_new, state, step0, n_steps, track, stream);
}

// Graphs of 1, 2, 4, ..., 2^(kGraphBits-1) pairs of steps: a run of n steps
// replays n / 2 pairs as whole 32-pair graphs and then the binary digits of
// the remainder, so a replay costs few graph launches (each costs several
// microseconds on the host and leaves a gap of ~9 us between launches on the
// device, as much as a fifth of a small grid's step).
constexpr int kGraphBits = 6;
constexpr int32_t kGraphLongPairs = 1 << (kGraphBits - 1);

// Captures `pairs` pairs of steps (u -> u_new -> u) into an executable graph
// and uploads it, so its first launch costs no more than the next ones.
// Each kernel reads its step from one device word and writes step + 1 to the
// other, so a replay continues from whatever step the words hold.
static int capture_pairs(fw_plan* plan, void* u, void* u_new, const StatePtrs& st,
                         int32_t* act, int32_t act_on, double act_thr, int32_t pairs,
                         hipGraphExec_t* out) {
  StepLauncher L = step_launcher(plan->desc.precision, plan->desc.model);
  hipStream_t cap;
  if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
    set_error("fw_run_graph: stream creation failed");
    return FW_E_HIP;
  }
  int rc = FW_OK;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    TrackArgs tr{};
    tr.act = act;
    tr.act_thr = act_thr;
    tr.act_on = act_on;
    tr.probe_row = -1;
    tr.stim_end_in = plan->step_words + 2;
    for (int32_t q = 0; q < pairs && !rc; ++q) {
      tr.step_in = plan->step_words;
      tr.step_out = plan->step_words + 1;
      rc = L(plan, u, u_new, st, tr, cap);
      tr.step_in = plan->step_words + 1;
      tr.step_out = plan->step_words;
      if (!rc) rc = L(plan, u_new, u, st, tr, cap);
    }
    e = hipStreamEndCapture(cap, &g);
  }
  hipGraphExec_t ex = nullptr;
  if (e == hipSuccess && !rc && g) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e == hipSuccess && !rc && ex) e = hipGraphUpload(ex, cap);
  if (e == hipSuccess && !rc && ex) e = hipStreamSynchronize(cap);
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess || rc || !ex) {
    if (ex) (void)hipGraphExecDestroy(ex);
    set_error("fw_run_graph: capture failed (%s)", hipGetErrorString(e));
    return rc ? rc : FW_E_HIP;
  }
  *out = ex;
  return FW_OK;
}

// The graph set fw_run_graph replays for these buffers, tracker and schedule
// (the plan's schedule is in every captured kernel's arguments; changing it
// drops the graphs), with the graphs an n_steps run needs captured.
static int ensure_graphs(fw_plan* plan, void* u, void* u_new, void* const* state,
                         int32_t n_steps, const fw_track_desc* track, int32_t* act_on_out,
                         fw_plan::GraphSet** set_out) {
  int rc = check_run_args(plan, u, u_new, state, n_steps, track);
  if (rc) return rc;
  if (track && track->n_probe > 0) {
    set_error("fw_run_graph: probes need per-step rows; use fw_run");
    return FW_E_UNSUPPORTED;
  }
  int32_t act_on = 0;
  int32_t* act = nullptr;
  if (track && track->act) {
    act = track->act;
    act_on = track->act_flags ? track->act_flags[0] : 1;
    if (track->act_flags)
      for (int32_t s = 1; s < n_steps; ++s)
        if (!!track->act_flags[s] != !!act_on) {
          set_error("fw_run_graph: the tracker gate must not change within the run");
          return FW_E_UNSUPPORTED;
        }
  }
  *act_on_out = act_on;
  const double act_thr = track ? track->act_thr : 0.0;
  const int ns = model_n_state(plan->desc.model);
  // graphs captured for another tracker or state are of no use
  bool same = plan->graph_keyed && plan->graph_act == act && plan->graph_act_on == act_on &&
              (!act || plan->graph_act_thr == act_thr);
  for (int k = 0; same && k < ns; ++k) same = plan->graph_state[k] == state[k];
  if (!same) {
    drop_graphs(plan);
    plan->graph_act = act;
    plan->graph_act_on = act_on;
    plan->graph_act_thr = act_thr;
    for (int k = 0; k < FW_MAX_STATE; ++k) plan->graph_state[k] = k < ns ? state[k] : nullptr;
    plan->graph_keyed = 1;
  }
  // the orientation's set, or the least recently used one
  fw_plan::GraphSet* gs = nullptr;
  for (auto& cand : plan->graphs)
    if (cand.u == u && cand.un == u_new) gs = &cand;
  if (!gs) {
    gs = plan->graphs[0].last_use <= plan->graphs[1].last_use ? &plan->graphs[0]
                                                               : &plan->graphs[1];
    for (auto& g : gs->g) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
    gs->u = u;
    gs->un = u_new;
  }
  gs->last_use = ++plan->graph_clock;
  const int32_t pairs = n_steps / 2;
  const StatePtrs st = make_state(state, ns);
  for (int bit = 0; bit < kGraphBits; ++bit) {
    const int32_t size = 1 << bit;
    const bool needed = bit == kGraphBits - 1 ? pairs >= size : ((pairs % kGraphLongPairs) & size);
    if (needed && !gs->g[bit]) {
      rc = capture_pairs(plan, u, u_new, st, act, act_on, act_thr, size, &gs->g[bit]);
      if (rc) return rc;
    }
  }
  *set_out = gs;
  return FW_OK;
}

int fw_run_graph_prepare(fw_plan* plan, void* u, void* u_new, void* const* state,
                         int32_t n_steps, const fw_track_desc* track) {
  int32_t act_on = 0;
  fw_plan::GraphSet* gs = nullptr;
  return ensure_graphs(plan, u, u_new, state, n_steps, track, &act_on, &gs);
}

int fw_run_graph(fw_plan* plan, void* u, void* u_new, void* const* state, int64_t step0,
                 int32_t n_steps, const fw_track_desc* track, void* stream) {
  int32_t act_on = 0;
  fw_plan::GraphSet* gs = nullptr;
  int rc = ensure_graphs(plan, u, u_new, state, n_steps, track, &act_on, &gs);
  if (rc) return rc;
  const int32_t pairs = n_steps / 2;
  hipStream_t s = as_stream(stream);
  if (pairs > 0) {
    hipLaunchKernelGGL(k_set_run_words, dim3(1), dim3(1), 0, s, plan->step_words,
                       (int32_t)step0, track ? clamp_step(track->stim_end) : 0);
    auto launch = [&](hipGraphExec_t g) -> int {
      const hipError_t e = hipGraphLaunch(g, s);
      if (e != hipSuccess) {
        set_error("fw_run_graph: launch failed: %s", hipGetErrorString(e));
        return FW_E_HIP;
      }
      return FW_OK;
    };
    for (int32_t q = 0; q + kGraphLongPairs <= pairs; q += kGraphLongPairs)
      if ((rc = launch(gs->g[kGraphBits - 1]))) return rc;
    const int32_t rest = pairs % kGraphLongPairs;
    for (int bit = kGraphBits - 2; bit >= 0; --bit)
      if ((rest >> bit) & 1)
        if ((rc = launch(gs->g[bit]))) return rc;
  }
  if (n_steps % 2) {
    // the last odd step runs as a plain launch on the original orientation
    fw_track_desc one{};
    uint8_t flag = (uint8_t)act_on;
    if (track) {
      one = *track;
      one.act_flags = &flag;
    }
    return fw_run(plan, u, u_new, state, step0 + 2 * pairs, 1, track ? &one : nullptr, stream);
  }
  return FW_OK;
}

int fw_stim_apply(fw_plan* plan, void* u, const fw_stim_desc* stim, void* stream) {
  if (!plan || !u || !stim || (stim->kind != 0 && stim->kind != 1)) {
    set_error("fw_stim_apply: bad arguments");
    return FW_E_INVALID;
  }
  return launch_stim(u_is_f64(plan->desc.model, plan->desc.precision), u, plan->desc.code, stim,
                     plan->desc.shape,
                     as_stream(stream));
}

int fw_plan_set_stims(fw_plan* plan, const fw_stim_sched* stims, int32_t n) {
  if (!plan || n < 0 || n > FW_MAX_STIMS || (n > 0 && !stims)) {
    set_error("fw_plan_set_stims: bad arguments (n = %d, at most %d)", n, FW_MAX_STIMS);
    return FW_E_INVALID;
  }
  const int nd = plan->desc.ndim;
  for (int32_t q = 0; q < n; ++q) {
    const fw_stim_sched& ss = stims[q];
    if (ss.stim.kind != 0 && ss.stim.kind != 1) {
      set_error("fw_plan_set_stims: stimulus %d: kind %d", q, ss.stim.kind);
      return FW_E_INVALID;
    }
    if (ss.step_lo < 0 || ss.step_hi < ss.step_lo || ss.step_hi > INT32_MAX) {
      set_error("fw_plan_set_stims: stimulus %d: steps [%lld, %lld]", q,
                (long long)ss.step_lo, (long long)ss.step_hi);
      return FW_E_INVALID;
    }
    for (int a = 0; a < 3; ++a) {
      const int64_t n_a = a < nd ? plan->desc.shape[a] : 1;
      if (ss.stim.lo[a] < 0 || ss.stim.hi[a] > n_a || ss.stim.lo[a] > ss.stim.hi[a]) {
        set_error("fw_plan_set_stims: stimulus %d: axis %d box [%lld, %lld) outside [0, %lld)",
                  q, a, (long long)ss.stim.lo[a], (long long)ss.stim.hi[a], (long long)n_a);
        return FW_E_INVALID;
      }
    }
    if (ss.values && ss.stim.kind != 0) {
      set_error("fw_plan_set_stims: stimulus %d: a value list needs a voltage stimulus", q);
      return FW_E_INVALID;
    }
  }
  const bool same = n == plan->n_stims &&
                    (n == 0 || memcmp(plan->stims, stims, n * sizeof(fw_stim_sched)) == 0);
  if (!same) {
    // captured graphs hold the old schedule in their kernel arguments
    drop_graphs(plan);
    plan->n_stims = n;
    if (n > 0) memcpy(plan->stims, stims, n * sizeof(fw_stim_sched));
  }
  return FW_OK;
}

int fw_stim_apply_step(fw_plan* plan, void* u, int64_t step, void* stream) {
  if (!plan || !u) {
    set_error("fw_stim_apply_step: bad arguments");
    return FW_E_INVALID;
  }
  for (int32_t q = 0; q < plan->n_stims; ++q) {
    const fw_stim_sched& ss = plan->stims[q];
    if (step < ss.step_lo || step > ss.step_hi) continue;
    const int rc = launch_stim(u_is_f64(plan->desc.model, plan->desc.precision), u,
                               plan->desc.code, &ss.stim, plan->desc.shape, as_stream(stream),
                               ss.values ? ss.values + (step - ss.step_lo) : nullptr);
    if (rc) return rc;
  }
  return FW_OK;
}

int fw_act_time_update(fw_plan* plan, int32_t* act, const void* u, double thr, int32_t step,
                       void* stream) {
  if (!plan || !act || !u) {
    set_error("fw_act_time_update: bad arguments");
    return FW_E_INVALID;
  }
  return launch_act(u_is_f64(plan->desc.model, plan->desc.precision), act, u, plan->N, thr, step,
                    as_stream(stream));
}

int fw_lat_update(fw_plan* plan, const void* u, int32_t* state, int32_t* steps,
                  int32_t capacity, double thr, int32_t step, int32_t* overflow, void* stream) {
  if (!plan || !u || !state || !steps || capacity <= 0 || !overflow) {
    set_error("fw_lat_update: bad arguments");
    return FW_E_INVALID;
  }
  return launch_lat(u_is_f64(plan->desc.model, plan->desc.precision), u, state, steps, capacity,
                    thr, step, plan->N, overflow, as_stream(stream));
}

int fw_fibrosis(int8_t* mesh, int32_t ndim, const int64_t* shape, const int64_t* lo,
                const int64_t* hi, const int64_t* block, double density, uint64_t seed,
                void* stream) {
  if (!mesh || !shape || !lo || !hi || (ndim != 2 && ndim != 3)) {
    set_error("fw_fibrosis: bad arguments");
    return FW_E_INVALID;
  }
  int64_t n[3] = {1, 1, 1}, l[3] = {0, 0, 0}, h[3] = {1, 1, 1}, b[3] = {1, 1, 1};
  for (int a = 0; a < ndim; ++a) {
    n[a] = shape[a];
    l[a] = lo[a];
    h[a] = hi[a];
    if (block) b[a] = block[a];
    if (n[a] < 1 || l[a] < 0 || l[a] > h[a] || h[a] > n[a] || b[a] < 1) {
      set_error("fw_fibrosis: axis %d: shape %lld, region [%lld, %lld), block %lld", a,
                (long long)n[a], (long long)l[a], (long long)h[a], (long long)b[a]);
      return FW_E_INVALID;
    }
  }
  const int64_t r[3] = {h[0] - l[0], h[1] - l[1], h[2] - l[2]};
  const int64_t count = r[0] * r[1] * r[2];
  if (count == 0) return FW_OK;
  const int64_t nb1 = (r[1] + b[1] - 1) / b[1], nb2 = (r[2] + b[2] - 1) / b[2];
  const int64_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_fibrosis, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256),
                     0, as_stream(stream), mesh, n[1], n[2], l[0], l[1], l[2], r[1], r[2],
                     count, b[0], b[1], b[2], nb1, nb2, block ? 0 : 1, density,
                     host_splitmix64(seed));
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("fw_fibrosis: launch failed: %s", hipGetErrorString(e));
    return FW_E_HIP;
  }
  return FW_OK;
}

int fw_spiral_tips(fw_plan* plan, const void* u_prev, const void* u, double thr, double* tips,
                   int32_t capacity, int32_t* count, void* stream) {
  if (!plan || !u_prev || !u || !tips || capacity <= 0 || !count) {
    set_error("fw_spiral_tips: bad arguments");
    return FW_E_INVALID;
  }
  const int64_t shape[3] = {plan->desc.shape[0], plan->desc.shape[1],
                            plan->desc.ndim == 3 ? plan->desc.shape[2] : 1};
  return launch_tips(u_is_f64(plan->desc.model, plan->desc.precision), u_prev, u, shape, thr,
                     tips, capacity, count, as_stream(stream));
}

int fw_ecg(fw_plan* plan, const void* u, const double* coords, int32_t n_points, double dr,
           double* out, void* stream) {
  if (!plan || !u || !coords || n_points <= 0 || !out || !(dr > 0)) {
    set_error("fw_ecg: bad arguments");
    return FW_E_INVALID;
  }
  EcgLauncher L = ecg_launcher(plan->desc.precision, plan->desc.model);
  if (!L) {
    set_error("fw_ecg: no kernel for this plan");
    return FW_E_UNSUPPORTED;
  }
  return L(plan, u, coords, n_points, dr, out, as_stream(stream));
}

}  // extern "C"

// ---------------------------------------------------------------- batches
namespace fw {

#define FW_DECL_BLAUNCHER(p, m) \
  int fw_batch_launch_p##p##_m##m(const BatchLaunch&, hipStream_t);
FW_DECL_BLAUNCHER(0, 0) FW_DECL_BLAUNCHER(0, 1) FW_DECL_BLAUNCHER(0, 2) FW_DECL_BLAUNCHER(0, 3)
FW_DECL_BLAUNCHER(0, 4) FW_DECL_BLAUNCHER(0, 5) FW_DECL_BLAUNCHER(0, 6) FW_DECL_BLAUNCHER(0, 7)
FW_DECL_BLAUNCHER(1, 0) FW_DECL_BLAUNCHER(1, 1) FW_DECL_BLAUNCHER(1, 2) FW_DECL_BLAUNCHER(1, 3)
FW_DECL_BLAUNCHER(1, 4) FW_DECL_BLAUNCHER(1, 5) FW_DECL_BLAUNCHER(1, 6) FW_DECL_BLAUNCHER(1, 7)

BatchLauncher batch_launcher(int precision, int model) {
  static const BatchLauncher table[2][8] = {
      {fw_batch_launch_p0_m0, fw_batch_launch_p0_m1, fw_batch_launch_p0_m2, fw_batch_launch_p0_m3,
       fw_batch_launch_p0_m4, fw_batch_launch_p0_m5, fw_batch_launch_p0_m6, fw_batch_launch_p0_m7},
      {fw_batch_launch_p1_m0, fw_batch_launch_p1_m1, fw_batch_launch_p1_m2, fw_batch_launch_p1_m3,
       fw_batch_launch_p1_m4, fw_batch_launch_p1_m5, fw_batch_launch_p1_m6, fw_batch_launch_p1_m7}};
  if (precision < 0 || precision > 1 || model < 0 || model > 7) return nullptr;
  return table[precision][model];
}

#define FW_DECL_TLAUNCHER(p, m) \
  int fw_batch_tb_launch_p##p##_m##m(const BatchLaunch&, const TbLaunch&, hipStream_t);
FW_DECL_TLAUNCHER(0, 0) FW_DECL_TLAUNCHER(0, 1) FW_DECL_TLAUNCHER(0, 2) FW_DECL_TLAUNCHER(0, 3)
FW_DECL_TLAUNCHER(0, 4) FW_DECL_TLAUNCHER(1, 0) FW_DECL_TLAUNCHER(1, 1) FW_DECL_TLAUNCHER(1, 2)
FW_DECL_TLAUNCHER(1, 3) FW_DECL_TLAUNCHER(1, 4)

BatchTbLauncher batch_tb_launcher(int precision, int model) {
  static const BatchTbLauncher table[2][5] = {
      {fw_batch_tb_launch_p0_m0, fw_batch_tb_launch_p0_m1, fw_batch_tb_launch_p0_m2,
       fw_batch_tb_launch_p0_m3, fw_batch_tb_launch_p0_m4},
      {fw_batch_tb_launch_p1_m0, fw_batch_tb_launch_p1_m1, fw_batch_tb_launch_p1_m2,
       fw_batch_tb_launch_p1_m3, fw_batch_tb_launch_p1_m4}};
  if (precision < 0 || precision > 1 || model < 0 || model > 4) return nullptr;
  return table[precision][model];
}

}  // namespace fw

static BatchLaunch batch_args(const fw_batch* b, int32_t step, const int32_t* in, int32_t* out) {
  BatchLaunch L;
  L.table = b->table;
  L.btis = b->btis;
  L.bstart = b->bstart;
  L.nblocks = (uint32_t)b->nblocks;
  L.per_xcd = (uint32_t)((b->nblocks + kXcds - 1) / kXcds);
  L.step_in = in;
  L.step_out = out;
  L.step = step;
  L.variant = b->variant & ((kBatchVariants - 1) | kBatchList);  // the single-step shape
  L.D = b->D;
  L.dt = b->dt;
  L.dr2 = b->dr2;
  L.inv_dr2 = b->inv_dr2;
  L.P = b->P;
  return L;
}

// The units of the multi-step shape (fw_batch_tb.hpp) for one tissue:
// column strips of 64 lanes whose own columns are K inside the window
// except at the tissue's edges, each split into runs of about `rows` rows.
static void tb_tissue_units(const fw_tissue_desc& d, int32_t slot, int K, int rows,
                            TbUnit* out, int64_t* n) {
  const int32_t ni = (int32_t)d.ni, nj = (int32_t)d.nj;
  const int32_t nseg = (ni + rows - 1) / rows;
  const int32_t len = (ni + nseg - 1) / nseg;
  for (int32_t c = 0; c < nj;) {
    const int32_t w0 = c == 0 ? 0 : c - K;
    const int32_t c1 = w0 + 64 >= nj ? nj : w0 + 64 - K;
    for (int32_t r0 = 0; r0 < ni; r0 += len) {
      const int32_t r1 = r0 + len < ni ? r0 + len : ni;
      if (out) out[*n] = TbUnit{slot, w0, c | (c1 << 16), r0 | (r1 << 16)};
      ++*n;
    }
    c = c1;
  }
}

static void drop_tb_graphs(fw_batch* b) {
  if (b->tb_graph) (void)hipGraphExecDestroy(b->tb_graph);
  if (b->tb_graph_long) (void)hipGraphExecDestroy(b->tb_graph_long);
  b->tb_graph = b->tb_graph_long = nullptr;
}

// Builds the unit table of the active tissues for the multi-step variant
// (the 5-point stencil's units first, then the 9-point stencil's).
static int build_units(fw_batch* b, const int32_t* slots, int32_t n_slots, hipStream_t s) {
  const int K = batch_tb_k(b->variant), rows = batch_tb_rows(b->variant);
  const int nst = model_n_state(b->model);
  int64_t n = 0;
  for (int32_t q = 0; q < n_slots; ++q) {
    const fw_tissue_desc& d = b->host[slots[q]];
    if (d.ni > 65535 || d.nj > 65535) {
      set_error("fw_batch_set_active: slot %d is %lld x %lld; the multi-step shapes take "
                "sides up to 65535", slots[q], (long long)d.ni, (long long)d.nj);
      return FW_E_INVALID;
    }
    for (int k = 0; k < nst; ++k)
      if (!d.state_alt[k]) {
        set_error("fw_batch_set_active: slot %d has no second state set (state_alt); the "
                  "multi-step shapes need one", slots[q]);
        return FW_E_INVALID;
      }
    tb_tissue_units(d, slots[q], K, rows, nullptr, &n);
  }
  if (n > INT32_MAX / 8) {
    set_error("fw_batch_set_active: too many units");
    return FW_E_INVALID;
  }
  TbUnit* hu = (TbUnit*)malloc(sizeof(TbUnit) * (size_t)(n > 0 ? n : 1));
  if (!hu) {
    set_error("fw_batch_set_active: out of host memory");
    return FW_E_NOMEM;
  }
  int64_t m = 0, n_iso = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (int32_t q = 0; q < n_slots; ++q) {
      const fw_tissue_desc& d = b->host[slots[q]];
      if ((d.stencil != FW_STENCIL_ISO) != (pass == 1)) continue;
      tb_tissue_units(d, slots[q], K, rows, hu, &m);
      if (pass == 0) n_iso = m;
    }
  if (n > b->cap_units) {
    if (hipStreamSynchronize(s) != hipSuccess) {
      free(hu);
      set_error("fw_batch_set_active: stream sync failed");
      return FW_E_HIP;
    }
    if (b->units) (void)hipFree(b->units);
    b->units = nullptr;
    b->cap_units = 0;
    drop_tb_graphs(b);  // they captured the old table's address
    const int64_t cap = n + n / 4 + 64;
    if (hipMalloc(&b->units, sizeof(TbUnit) * cap) != hipSuccess) {
      free(hu);
      b->units = nullptr;
      set_error("fw_batch_set_active: allocation failed");
      return FW_E_NOMEM;
    }
    b->cap_units = (int32_t)cap;
  }
  if (n && (hipMemcpyAsync(b->units, hu, sizeof(TbUnit) * n, hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)) {
    free(hu);
    set_error("fw_batch_set_active: copy failed");
    return FW_E_HIP;
  }
  free(hu);
  b->n_units = (int32_t)n;
  b->n_units_iso = (int32_t)n_iso;
  b->tb_variant = b->variant;
  return FW_OK;
}

extern "C" {

int fw_batch_create(int32_t model, int32_t precision, double D_model, double dt, double dr2,
                    const double* params, int32_t n_params, int32_t max_tissues, fw_batch** out) {
  if (!out || !params || model < 0 || model >= FW_N_MODELS || max_tissues <= 0 ||
      (precision != FW_PREC_FP64 && precision != FW_PREC_FP32) || !(dt > 0) || !(dr2 > 0)) {
    set_error("fw_batch_create: bad arguments");
    return FW_E_INVALID;
  }
  if (n_params != model_n_params(model)) {
    set_error("fw_batch_create: model %d takes %d parameters (got %d)", model,
              model_n_params(model), n_params);
    return FW_E_INVALID;
  }
  *out = nullptr;
  fw_batch* b = new (std::nothrow) fw_batch();
  if (!b) return FW_E_NOMEM;
  b->model = model;
  b->precision = precision;
  b->max_tissues = max_tissues;
  b->variant = batch_default_variant(model);
  b->tb_variant = -1;
  b->active = new (std::nothrow) int32_t[max_tissues];
  b->D = D_model;
  b->dt = dt;
  b->dr2 = dr2;
  int e2;
  const double mant = frexp(dr2, &e2);
  b->inv_dr2 = (mant == 0.5 && e2 > -1000 && e2 < 1000) ? ldexp(1.0, 1 - e2) : 0.0;
  for (int q = 0; q < FW_MAX_PARAMS; ++q) b->P.v[q] = q < n_params ? params[q] : 0.0;
  complete_params(model, b->P);
  b->host = new (std::nothrow) fw_tissue_desc[max_tissues]();
  b->lists = new (std::nothrow) int32_t*[max_tissues]();
  b->list_n = new (std::nothrow) int32_t[max_tissues]();
  b->list_cap = new (std::nothrow) int64_t[max_tissues]();
  if (!b->host || !b->active || !b->lists || !b->list_n || !b->list_cap ||
      hipMalloc(&b->table, sizeof(BatchTissue) * max_tissues) != hipSuccess ||
      hipMalloc(&b->step_words, 4 * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&b->flags, sizeof(int32_t) * max_tissues) != hipSuccess) {
    set_error("fw_batch_create: allocation failed");
    (void)fw_batch_destroy(b);
    return FW_E_NOMEM;
  }
  // cleared before any stream (blocking or not) can use the batch
  if (hipMemset(b->table, 0, sizeof(BatchTissue) * max_tissues) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    set_error("fw_batch_create: table init failed");
    (void)fw_batch_destroy(b);
    return FW_E_HIP;
  }
  *out = b;
  return FW_OK;
}

int fw_batch_destroy(fw_batch* b) {
  if (!b) return FW_OK;
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->graph_long) (void)hipGraphExecDestroy(b->graph_long);
  if (b->tb_graph) (void)hipGraphExecDestroy(b->tb_graph);
  if (b->tb_graph_long) (void)hipGraphExecDestroy(b->tb_graph_long);
  if (b->units) (void)hipFree(b->units);
  if (b->table) (void)hipFree(b->table);
  if (b->btis) (void)hipFree(b->btis);
  if (b->bstart) (void)hipFree(b->bstart);
  if (b->step_words) (void)hipFree(b->step_words);
  if (b->flags) (void)hipFree(b->flags);
  if (b->lists)
    for (int32_t q = 0; q < b->max_tissues; ++q)
      if (b->lists[q]) (void)hipFree(b->lists[q]);
  delete[] b->lists;
  delete[] b->list_n;
  delete[] b->list_cap;
  delete[] b->host;
  delete[] b->active;
  delete b;
  return FW_OK;
}

int fw_batch_set_tissue(fw_batch* b, int32_t slot, const fw_tissue_desc* d, void* stream) {
  if (!b || !d || slot < 0 || slot >= b->max_tissues) {
    set_error("fw_batch_set_tissue: bad arguments");
    return FW_E_INVALID;
  }
  if (d->n != d->ni * d->nj || d->ni < 3 || d->nj < 3 || d->n > INT32_MAX / 4 || !d->u ||
      !d->un || !d->code || (d->stencil != FW_STENCIL_ISO && d->stencil != FW_STENCIL_ANISO) ||
      (d->probe && d->probe_idx >= d->n)) {
    set_error("fw_batch_set_tissue: malformed descriptor for slot %d", slot);
    return FW_E_INVALID;
  }
  for (int k = 0; k < model_n_state(b->model); ++k)
    if (!d->state[k]) {
      set_error("fw_batch_set_tissue: state %d of slot %d is null", k, slot);
      return FW_E_INVALID;
    }
  hipStream_t s = as_stream(stream);
  // no updated node on the outer layer (the kernel then needs no column
  // guard) and no frozen node (non-updated nodes are written back unchanged)
  int32_t bad = 0;
  if (hipMemsetAsync(b->step_words + 2, 0, sizeof(int32_t), s) != hipSuccess) {
    set_error("fw_batch_set_tissue: flag reset failed");
    return FW_E_HIP;
  }
  const int64_t nblk = (d->n + 255) / 256;
  hipLaunchKernelGGL(k_check_codes, dim3((unsigned)(nblk < 4096 ? nblk : 4096)), dim3(256), 0,
                     s, d->code, d->ni, d->nj, b->step_words + 2);
  if (hipMemcpyAsync(&bad, b->step_words + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("fw_batch_set_tissue: border check failed");
    return FW_E_HIP;
  }
  if (bad & 1) {
    set_error("fw_batch_set_tissue: slot %d has updated nodes (code & 3 == 1) on its outer "
              "layer", slot);
    return FW_E_INVALID;
  }
  if (bad & 2) {
    set_error("fw_batch_set_tissue: slot %d has frozen nodes (code bit 1); batched tissues "
              "take no special boundaries", slot);
    return FW_E_INVALID;
  }
  // models 5-7: the node-list shapes' list of the tissue's updated nodes and
  // its probe node, ascending, built on the host from the codes (the stream
  // has drained: no launch still reads the slot's previous list)
  int32_t n_list = 0;
  if (b->model >= 5) {
    std::vector<uint8_t> codes((size_t)d->n);
    if (hipMemcpyAsync(codes.data(), d->code, (size_t)d->n, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("fw_batch_set_tissue: code copy failed");
      return FW_E_HIP;
    }
    const int64_t pidx = d->probe ? d->probe_idx : -1;
    std::vector<int32_t> list;
    list.reserve((size_t)d->n);
    for (int64_t c = 0; c < d->n; ++c)
      if ((codes[(size_t)c] & 3) == 1 || c == pidx) list.push_back((int32_t)c);
    n_list = (int32_t)list.size();
    if (n_list > b->list_cap[slot]) {
      if (b->lists[slot]) (void)hipFree(b->lists[slot]);
      b->lists[slot] = nullptr;
      b->list_cap[slot] = 0;
      if (hipMalloc(&b->lists[slot], sizeof(int32_t) * (size_t)n_list) != hipSuccess) {
        set_error("fw_batch_set_tissue: node list allocation failed");
        return FW_E_NOMEM;
      }
      b->list_cap[slot] = n_list;
    }
    if (n_list && (hipMemcpyAsync(b->lists[slot], list.data(), sizeof(int32_t) * (size_t)n_list,
                                  hipMemcpyHostToDevice, s) != hipSuccess ||
                   hipStreamSynchronize(s) != hipSuccess)) {
      set_error("fw_batch_set_tissue: node list copy failed");
      return FW_E_HIP;
    }
  }
  BatchTissue r;
  memset(&r, 0, sizeof(r));
  r.nodes = b->model >= 5 ? b->lists[slot] : nullptr;
  r.n_nodes = n_list;
  r.u = d->u;
  r.un = d->un;
  r.code = d->code;
  r.act = d->act;
  r.probe = d->probe;
  for (int k = 0; k < FW_MAX_STATE; ++k) r.state[k] = d->state[k];
  for (int k = 0; k < FW_BATCH_ALT_STATE; ++k) r.state_alt[k] = d->state_alt[k];
  for (int q = 0; q < 4; ++q) {
    r.dd[q] = d->d[q];
    r.df[q] = (float)d->d[q];
  }
  r.act_thr = d->act_thr;
  r.act_thr_f = (float)d->act_thr;
  r.n = (int32_t)d->n;
  r.ni = (int32_t)d->ni;
  r.nj = (int32_t)d->nj;
  r.step0 = d->step0;
  r.act_start = d->act_start;
  r.probe_idx = d->probe ? d->probe_idx : -1;
  r.probe_len = d->probe ? d->probe_len : 0;
  r.stencil = d->stencil;
  if (hipMemcpyAsync(static_cast<BatchTissue*>(b->table) + slot, &r, sizeof(r),
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("fw_batch_set_tissue: copy failed");
    return FW_E_HIP;
  }
  b->host[slot] = *d;
  b->list_n[slot] = n_list;
  return FW_OK;
}

int fw_batch_release(fw_batch* b, int32_t slot) {
  if (!b || slot < 0 || slot >= b->max_tissues) {
    set_error("fw_batch_release: bad arguments");
    return FW_E_INVALID;
  }
  b->host[slot] = fw_tissue_desc();
  return FW_OK;
}

int fw_batch_block_nodes(const fw_batch* b) {
  if (!b) {
    set_error("fw_batch_block_nodes: bad arguments");
    return FW_E_INVALID;
  }
  return batch_unroll(b->variant) * kBlock;
}

int fw_batch_variants(int32_t model, int32_t* variants, int32_t capacity) {
  if (model < 0 || model >= FW_N_MODELS || (capacity > 0 && !variants)) {
    set_error("fw_batch_variants: bad arguments");
    return FW_E_INVALID;
  }
  int32_t n = 0;
  auto add = [&](int32_t v) {
    if (n < capacity) variants[n] = v;
    ++n;
  };
  for (int32_t v = 0; v < kBatchVariants; ++v)
    if (batch_variant_ok(model, v)) add(v);
  // node-list shapes over each single-step shape (models 5-7)
  for (int32_t v = 0; v < kBatchVariants; ++v)
    if (batch_variant_ok(model, v | kBatchList)) add(v | kBatchList);
  // multi-step shapes over the XCD-contiguous 4-node single-step shape, with
  // and without the cache hint
  for (int32_t base : {kBatchVec | kBatchXcd | kBatchNT, kBatchVec | kBatchXcd})
    for (int32_t k = 0; k < kBatchTbKs; ++k)
      for (int32_t r = 0; r < kBatchTbRowSizes; ++r) {
        const int32_t v = batch_tb_variant(base, k, r);
        if (batch_variant_ok(model, v)) add(v);
      }
  return n;
}

int fw_batch_get_variant(const fw_batch* b) {
  if (!b) {
    set_error("fw_batch_get_variant: bad arguments");
    return FW_E_INVALID;
  }
  return b->variant;
}

int fw_batch_set_variant(fw_batch* b, int32_t variant, void* stream) {
  if (!b || !batch_variant_ok(b->model, variant)) {
    set_error("fw_batch_set_variant: variant %d not available for this batch", variant);
    return FW_E_INVALID;
  }
  if (variant == b->variant) return FW_OK;
  const int32_t old = b->variant;
  b->variant = variant;
  if (b->n_active > 0 && (batch_unroll(old) != batch_unroll(variant) ||
                          ((old | variant) & kBatchTB) || ((old ^ variant) & kBatchList))) {
    // the block table's granularity changes with the nodes per work-item and
    // its extent with the node lists, the unit table with the multi-step shape
    int rc = fw_batch_set_active(b, b->active, b->n_active, stream);
    if (rc) {
      b->variant = old;
      (void)fw_batch_set_active(b, b->active, b->n_active, stream);
      return rc;
    }
  }
  return FW_OK;
}

int fw_batch_set_active(fw_batch* b, const int32_t* slots, int32_t n_slots, void* stream) {
  if (!b || n_slots < 0 || n_slots > b->max_tissues || (n_slots && !slots)) {
    set_error("fw_batch_set_active: bad arguments");
    return FW_E_INVALID;
  }
  const int64_t per = (int64_t)batch_unroll(b->variant) * kBlock;
  const bool listed = (b->variant & kBatchList) != 0;
  // blocks cover a tissue's nodes, or its node list under the list shapes
  auto extent = [&](int32_t s) -> int64_t { return listed ? b->list_n[s] : b->host[s].n; };
  int64_t nblocks = 0;
  for (int32_t q = 0; q < n_slots; ++q) {
    const int32_t s = slots[q];
    if (s < 0 || s >= b->max_tissues || !b->host[s].u) {
      set_error("fw_batch_set_active: slot %d is not occupied", s);
      return FW_E_INVALID;
    }
    nblocks += (extent(s) + per - 1) / per;
  }
  if (nblocks > INT32_MAX / 2) {
    set_error("fw_batch_set_active: too many blocks");
    return FW_E_INVALID;
  }
  int32_t* hb = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)(nblocks > 0 ? nblocks : 1));
  if (!hb) {
    set_error("fw_batch_set_active: out of host memory");
    return FW_E_NOMEM;
  }
  int32_t* ht = hb;
  int32_t* hs = hb + nblocks;
  int64_t q0 = 0;
  for (int32_t q = 0; q < n_slots; ++q) {
    const int64_t n = extent(slots[q]);
    for (int64_t st = 0; st < n; st += per, ++q0) {
      ht[q0] = slots[q];
      hs[q0] = (int32_t)st;
    }
  }
  hipStream_t s = as_stream(stream);
  if (nblocks > b->cap_blocks) {
    if (hipStreamSynchronize(s) != hipSuccess) {
      free(hb);
      set_error("fw_batch_set_active: stream sync failed");
      return FW_E_HIP;
    }
    if (b->btis) (void)hipFree(b->btis);
    if (b->bstart) (void)hipFree(b->bstart);
    b->btis = b->bstart = nullptr;
    const int64_t cap = nblocks + nblocks / 4 + 64;
    if (hipMalloc(&b->btis, sizeof(int32_t) * cap) != hipSuccess ||
        hipMalloc(&b->bstart, sizeof(int32_t) * cap) != hipSuccess) {
      free(hb);
      set_error("fw_batch_set_active: allocation failed");
      b->cap_blocks = 0;
      b->nblocks = 0;
      return FW_E_NOMEM;
    }
    b->cap_blocks = (int32_t)cap;
    // the graphs captured the old table's addresses
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graph_long) (void)hipGraphExecDestroy(b->graph_long);
    b->graph = b->graph_long = nullptr;
  }
  if (nblocks &&
      (hipMemcpyAsync(b->btis, ht, sizeof(int32_t) * nblocks, hipMemcpyHostToDevice, s) !=
           hipSuccess ||
       hipMemcpyAsync(b->bstart, hs, sizeof(int32_t) * nblocks, hipMemcpyHostToDevice, s) !=
           hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess)) {
    free(hb);
    set_error("fw_batch_set_active: copy failed");
    return FW_E_HIP;
  }
  free(hb);
  b->nblocks = (int32_t)nblocks;
  b->tb_variant = -1;
  if (b->variant & kBatchTB) {
    const int rc = build_units(b, slots, n_slots, s);
    if (rc) return rc;
  }
  if (slots != b->active)
    for (int32_t q = 0; q < n_slots; ++q) b->active[q] = slots[q];
  b->n_active = n_slots;
  return FW_OK;
}

// Captures `pairs` pairs of batched steps into an executable graph; the step
// words carry the step from one launch to the next, as for single tissues.
static int capture_batch(fw_batch* b, BatchLauncher L, int32_t pairs, hipGraphExec_t* out) {
  hipStream_t cap;
  if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
    set_error("fw_batch_run: stream creation failed");
    return FW_E_HIP;
  }
  hipGraph_t g = nullptr;
  int rc = FW_OK;
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int32_t q = 0; q < pairs && !rc; ++q) {
      rc = L(batch_args(b, 0, b->step_words, b->step_words + 1), cap);
      if (!rc) rc = L(batch_args(b, 0, b->step_words + 1, b->step_words), cap);
    }
    e = hipStreamEndCapture(cap, &g);
  }
  hipGraphExec_t ex = nullptr;
  if (e == hipSuccess && !rc && g) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess || rc || !ex) {
    set_error("fw_batch_run: capture failed (%s)", hipGetErrorString(e));
    return rc ? rc : FW_E_HIP;
  }
  *out = ex;
  return FW_OK;
}

}  // extern "C"

// n single steps of the batch's single-step shape from step0
static int batch_run_single(fw_batch* b, int64_t step0, int32_t n_steps, int32_t use_graph,
                            hipStream_t s) {
  BatchLauncher L = batch_launcher(b->precision, b->model);
  int32_t done = 0;
  if (use_graph && n_steps >= 2) {
    const int32_t pairs = n_steps / 2;
    if (b->graph_nblocks != b->nblocks || b->graph_variant != b->variant) {
      // captured for another block table or launch shape
      if (b->graph) (void)hipGraphExecDestroy(b->graph);
      if (b->graph_long) (void)hipGraphExecDestroy(b->graph_long);
      b->graph = b->graph_long = nullptr;
      b->graph_nblocks = b->nblocks;
      b->graph_variant = b->variant;
    }
    if (!b->graph) {
      int rc = capture_batch(b, L, 1, &b->graph);
      if (rc) return rc;
    }
    if (pairs >= kGraphLongPairs && !b->graph_long) {
      int rc = capture_batch(b, L, kGraphLongPairs, &b->graph_long);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, s, b->step_words, (int32_t)step0);
    int32_t q = 0;
    for (; q + kGraphLongPairs <= pairs; q += kGraphLongPairs)
      if (hipGraphLaunch(b->graph_long, s) != hipSuccess) {
        set_error("fw_batch_run: graph launch failed");
        return FW_E_HIP;
      }
    for (; q < pairs; ++q)
      if (hipGraphLaunch(b->graph, s) != hipSuccess) {
        set_error("fw_batch_run: graph launch failed");
        return FW_E_HIP;
      }
    done = 2 * pairs;
  }
  for (int32_t q = done; q < n_steps; ++q) {
    int rc = L(batch_args(b, (int32_t)(step0 + q), nullptr, nullptr), s);
    if (rc) return rc;
  }
  return FW_OK;
}

// One pair of multi-step launches (phase 0 then 1): 2K steps from `step`,
// or from the step word `in` (the pair ends with it again) in a graph.
static int tb_pair(fw_batch* b, int32_t step, int32_t* in, int32_t* mid, hipStream_t s) {
  BatchTbLauncher T = batch_tb_launcher(b->precision, b->model);
  if (!T) {
    set_error("fw_batch_run: model %d has no multi-step shape", b->model);
    return FW_E_UNSUPPORTED;
  }
  const int K = batch_tb_k(b->variant);
  const int32_t n_an = b->n_units - b->n_units_iso;
  for (int phase = 0; phase < 2; ++phase) {
    BatchLaunch L = batch_args(b, step + phase * K, phase ? mid : in, phase ? in : mid);
    L.variant = b->variant;
    for (int an = 0; an < 2; ++an) {
      const int32_t nu = an ? n_an : b->n_units_iso;
      if (nu == 0) continue;
      TbLaunch t;
      t.units = static_cast<const TbUnit*>(b->units) + (an ? b->n_units_iso : 0);
      t.n_units = (uint32_t)nu;
      t.nblocks = (uint32_t)((nu + kTbWaves - 1) / kTbWaves);
      t.per_xcd = (t.nblocks + kXcds - 1) / kXcds;
      t.phase = phase;
      t.aniso = an;
      const int rc = T(L, t, s);
      if (rc) return rc;
    }
  }
  return FW_OK;
}

static int capture_tb(fw_batch* b, int32_t pairs, hipGraphExec_t* out) {
  hipStream_t cap;
  if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
    set_error("fw_batch_run: stream creation failed");
    return FW_E_HIP;
  }
  hipGraph_t g = nullptr;
  int rc = FW_OK;
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int32_t q = 0; q < pairs && !rc; ++q)
      rc = tb_pair(b, 0, b->step_words, b->step_words + 1, cap);
    e = hipStreamEndCapture(cap, &g);
  }
  hipGraphExec_t ex = nullptr;
  if (e == hipSuccess && !rc && g) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e == hipSuccess && !rc && ex) e = hipGraphUpload(ex, cap);
  if (e == hipSuccess && !rc) e = hipStreamSynchronize(cap);
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess || rc || !ex) {
    if (ex) (void)hipGraphExecDestroy(ex);
    set_error("fw_batch_run: capture failed (%s)", hipGetErrorString(e));
    return rc ? rc : FW_E_HIP;
  }
  *out = ex;
  return FW_OK;
}

// multi-step pairs per long graph
constexpr int32_t kBatchTbLongPairs = 8;

extern "C" {

int fw_batch_run(fw_batch* b, int64_t step0, int32_t n_steps, int32_t use_graph, void* stream) {
  if (!b || n_steps < 0) {
    set_error("fw_batch_run: bad arguments");
    return FW_E_INVALID;
  }
  if (b->nblocks == 0 || n_steps == 0) return FW_OK;
  hipStream_t s = as_stream(stream);
  if (!(b->variant & kBatchTB)) return batch_run_single(b, step0, n_steps, use_graph, s);
  if (b->tb_variant != b->variant) {
    set_error("fw_batch_run: the unit table was not built for this launch shape");
    return FW_E_INVALID;
  }
  // single steps until every active tissue has started
  int64_t lead = 0;
  for (int32_t q = 0; q < b->n_active; ++q) {
    const int64_t d = (int64_t)b->host[b->active[q]].step0 - step0;
    if (d > lead) lead = d;
  }
  if (lead > n_steps) lead = n_steps;
  if (lead > 0) {
    const int rc = batch_run_single(b, step0, (int32_t)lead, use_graph, s);
    if (rc) return rc;
  }
  const int32_t K2 = 2 * batch_tb_k(b->variant);
  const int64_t t0 = step0 + lead;
  const int32_t pairs = (int32_t)((n_steps - lead) / K2);
  if (pairs > 0 && use_graph) {
    if (b->tb_graph_units != b->n_units || b->tb_graph_iso != b->n_units_iso ||
        b->tb_graph_variant != b->variant) {
      drop_tb_graphs(b);
      b->tb_graph_units = b->n_units;
      b->tb_graph_iso = b->n_units_iso;
      b->tb_graph_variant = b->variant;
    }
    if (!b->tb_graph) {
      const int rc = capture_tb(b, 1, &b->tb_graph);
      if (rc) return rc;
    }
    if (pairs >= kBatchTbLongPairs && !b->tb_graph_long) {
      const int rc = capture_tb(b, kBatchTbLongPairs, &b->tb_graph_long);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, s, b->step_words, (int32_t)t0);
    int32_t q = 0;
    for (; q + kBatchTbLongPairs <= pairs; q += kBatchTbLongPairs)
      if (hipGraphLaunch(b->tb_graph_long, s) != hipSuccess) {
        set_error("fw_batch_run: graph launch failed");
        return FW_E_HIP;
      }
    for (; q < pairs; ++q)
      if (hipGraphLaunch(b->tb_graph, s) != hipSuccess) {
        set_error("fw_batch_run: graph launch failed");
        return FW_E_HIP;
      }
  } else {
    for (int32_t q = 0; q < pairs; ++q) {
      const int rc = tb_pair(b, (int32_t)(t0 + (int64_t)q * K2), nullptr, nullptr, s);
      if (rc) return rc;
    }
  }
  const int64_t done = lead + (int64_t)pairs * K2;
  if (done < n_steps)
    return batch_run_single(b, step0 + done, (int32_t)(n_steps - done), use_graph, s);
  return FW_OK;
}

int fw_batch_nonfinite(fw_batch* b, int64_t step, int32_t* flags, void* stream) {
  if (!b || !flags) {
    set_error("fw_batch_nonfinite: bad arguments");
    return FW_E_INVALID;
  }
  hipStream_t s = as_stream(stream);
  const int32_t nt = b->max_tissues;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, b->flags,
                     nt, 0);
  if (b->nblocks > 0) {
    const int32_t per = batch_unroll(b->variant) * kBlock;
    if (u_is_f64(b->model, b->precision))
      hipLaunchKernelGGL(k_batch_nonfinite<double>, dim3((unsigned)b->nblocks), dim3(256), 0, s,
                         static_cast<const BatchTissue*>(b->table), b->btis, b->bstart, per,
                         (int32_t)step, b->flags);
    else
      hipLaunchKernelGGL(k_batch_nonfinite<float>, dim3((unsigned)b->nblocks), dim3(256), 0, s,
                         static_cast<const BatchTissue*>(b->table), b->btis, b->bstart, per,
                         (int32_t)step, b->flags);
  }
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(flags, b->flags, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, s) !=
          hipSuccess) {
    set_error("fw_batch_nonfinite: check failed");
    return FW_E_HIP;
  }
  return FW_OK;
}

int fw_batch_stim(fw_batch* b, int32_t slot, int64_t step, const fw_stim_desc* stim,
                  void* stream) {
  if (!b || !stim || slot < 0 || slot >= b->max_tissues || (stim->kind != 0 && stim->kind != 1)) {
    set_error("fw_batch_stim: bad arguments");
    return FW_E_INVALID;
  }
  const fw_tissue_desc& d = b->host[slot];
  if (!d.u) {
    set_error("fw_batch_stim: slot %d is empty", slot);
    return FW_E_INVALID;
  }
  const int64_t ls = step - d.step0;
  void* u = (ls & 1) ? d.un : d.u;
  const int64_t shape[3] = {d.ni, d.nj, 1};
  return launch_stim(u_is_f64(b->model, b->precision), u, d.code, stim, shape, as_stream(stream));
}

}  // extern "C"
