// fw_models.hpp -- per-cell ionic updates for gfx950, templated on the
// storage/arithmetic type R (double for the FP64 path, float for the FP32
// path).  Calcium and ion concentrations of TP06 are always double.
//
// Each update reads the node's pre-step potential `u` and its state, writes
// the state in place and returns the new u_new given the diffusion result
// `un` -- the reference's "ionic_kernel after diffusion_kernel" order.
// Operation order follows the reference statements (paths relative to the
// finitewave source); in the FP64 translation unit, built with
// -ffp-contract=off, the arithmetic is the reference's bit for bit except
// for the libm implementation of exp/log/tanh/pow.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "fw_math.hpp"

namespace fw {

constexpr int kBlock = 256;  // work-items per block of the node-parallel kernels
constexpr int kXcds = 8;     // accelerator complex dies of an MI355X

struct Params {
  double v[64];  // the model's parameters (then derived quantities), as given
  float f[64];   // the same rounded to float once on the host (complete_params)
};

// parameter q in the arithmetic type: the FP32 path reads the host's float
// copy, a kernel-argument value, instead of converting per work-item
template <typename R>
__device__ __forceinline__ R par(const Params& P, int q) {
  if constexpr (sizeof(R) == 4) return P.f[q];
  else return P.v[q];
}

struct StatePtrs {
  void* p[24];
};

// The same pointers, for kernels that stream the state: model code reading
// and writing through these gets non-temporal accesses (StateRef below).
struct StreamStatePtrs {
  void* p[24];
};

// ---- math dispatch --------------------------------------------------------
// FP64: the exp/log of fw_math.hpp, whose last bits follow the host libm the
// reference's numba kernels call (exp bit for bit); FP32: the device library.
__device__ __forceinline__ double fexp(double x) { return t128_exp(x); }
__device__ __forceinline__ float fexp(float x) { return __expf(x); }
__device__ __forceinline__ double flog(double x) { return dd_log(x); }
__device__ __forceinline__ float flog(float x) { return __logf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double ftanh(double x) { return tanh(x); }
__device__ __forceinline__ float ftanh(float x) { return tanhf(x); }
__device__ __forceinline__ double fpow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float fpow(float x, float y) { return powf(x, y); }

// (e^y - 1)/y, continuous at y = 0.  The FP32 path evaluates the rate
// expressions of the form x / (e^(k x) - 1) through it: in float e^(k x)
// rounds to exactly 1 within about 1e-7 of the removable singularity and the
// literal form divides by zero there.
__device__ __forceinline__ float exprel(float y) {
  return fabsf(y) < 1e-4f ? 1.0f + 0.5f * y : expm1f(y) / y;
}

// a / b for a literal b.  FP64 path: the correctly rounded quotient in five
// operations instead of a full division, from y = 1/b correctly rounded
// (folded at compile time): q0 = a y is within 1.5 ulp of a/b, one
// correction q1 = q0 + (a - b q0) y is faithful, and a second gives RN(a/b)
// by Markstein's theorem (y within half an ulp of 1/b, q1 within one ulp of
// a/b; the remainder a - b q1 is exact).  Holds while a/b is a normal number
// or zero (a zero quotient comes out +0).  FP32 path: the plain division.
template <typename R>
__device__ __forceinline__ R div_rn(R a, R b, R y) {  // y = 1/b correctly rounded
  const R q0 = a * y;
  const R q1 = fma(fma(-b, q0, a), y, q0);
  return fma(fma(-b, q1, a), y, q1);
}

template <typename R>
__device__ __forceinline__ R divc(R a, R b) {
  if constexpr (sizeof(R) == 8) return div_rn(a, b, R(1) / b);
  else return a / b;
}

// a / b for a node-dependent b.  FP64: the core of the IEEE division
// sequence, without its operand scaling (two v_div_scale), its scaled
// correction and its special-case fix-up: r = 1/b from the hardware estimate
// refined by two Newton steps, q = a r, then q + (a - b q) r.  These are the
// operations and roundings of the full sequence, so the quotient is the same
// bits wherever the full sequence would not rescale: a, b, 1/b and a/b normal
// and their exponents less than 768 apart (the models' rates, gates,
// concentrations and time constants stay far inside that while the state is
// finite).  It differs for b infinite (NaN instead of 0) and for a = -0
// (+0).  FP32: the plain division.
template <typename R>
__device__ __forceinline__ R dv(R a, R b) {
  if constexpr (sizeof(R) == 8) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
  } else {
    return a / b;
  }
}

// TP06 quantities that are the same for every node (functions of the
// parameters only), appended after its 49 parameters by complete_params on the
// host, in the operation order of the expressions they replace: IEEE +, *, /
// and sqrt round the same on host and device, so the FP64 path's values are
// unchanged while each node saves ten divisions and a square root.  The FP32
// path rounds the same double values to float once.
constexpr int kTp06Derived = 49;
enum Tp06Derived {
  kInvVcF,      // 1.0 / (Vc * F)
  kInvVssF2,    // 1.0 / (2 * Vss * F)
  kFFoverRT,    // F * F / (R * T)
  kRT,          // R * T
  kInvRT,       // 1.0 / (R * T)
  kInvNaCaK,    // 1.0 / (KmNai * KmNai * KmNai + nao * nao * nao)
  kInvKmCaCao,  // 1.0 / (KmCa + cao)
  kKoOverKoKmK, // ko / (ko + KmK)
  kSqrtKo54,    // sqrt(ko / 5.4)
  kVcOverVss,   // Vc / Vss
  kVsrOverVss,  // Vsr / Vss
  kTp06NDerived
};
static_assert(kTp06Derived + kTp06NDerived <= 64, "parameter block size");

// state arrays live in global memory: accesses through the global address
// space compile to global (not flat) loads and stores
template <typename T>
using global_t = __attribute__((address_space(1))) T;

template <typename T>
__device__ __forceinline__ global_t<T>& ref(const StatePtrs& S, int k, int64_t c) {
  return ((global_t<T>*)S.p[k])[c];
}

// A node's state value in global memory, accessed with the non-temporal hint.
// Each value is read once and written once per step, so when a step's data is
// much larger than the caches, streaming it past them reaches a higher share
// of HBM bandwidth; when it nearly fits the 256 MB Infinity Cache the default
// policy is faster (it keeps part of the state there between steps).  Model
// code binds a reference with auto&&, reads it into a local and assigns the
// result back, so the same code serves both kinds of pointers.
template <typename T>
struct StateRef {
  global_t<T>* p;
  __device__ __forceinline__ operator T() const { return __builtin_nontemporal_load(p); }
  __device__ __forceinline__ const StateRef& operator=(T v) const {
    __builtin_nontemporal_store(v, p);
    return *this;
  }
};

template <typename T>
__device__ __forceinline__ StateRef<T> ref(const StreamStatePtrs& S, int k, int64_t c) {
  return StateRef<T>{(global_t<T>*)S.p[k] + c};
}

template <bool NT, typename T>
__device__ __forceinline__ T ld_state(const global_t<T>* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <bool NT, typename T>
__device__ __forceinline__ void st_state(global_t<T>* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// The state of one node held in registers (models whose state is all of the
// arithmetic type): the same update code then reads and writes registers,
// so a kernel can load the state of several nodes before storing any.
template <typename R, int NS>
struct RegState {
  R v[NS > 0 ? NS : 1];
};

template <typename T, typename R, int NS>
__device__ __forceinline__ T& ref(RegState<R, NS>& S, int k, int64_t) {
  return S.v[k];
}

// Stored form of state k: its value minus state_offset(model, prec, k).  The
// FP32 path keeps TP06's nai, Ki (k = 3, 4) and sodium inactivation gates h,
// j (k = 6, 7) as float differences from their initial values
// (tp06_2d.py:155-159); every other state, and every state of the FP64 path,
// is stored as its value.  A float near 138.3 resolves 1.5e-5 and one near
// 0.75 resolves 6e-8; the differences stay small over a simulation (the
// concentrations drift slowly, the gates rest near 0.72-0.77), where their
// ulp is a few hundred to tens of thousands of times finer.
__host__ __device__ constexpr double state_offset(int model, int prec, int k) {
  return (prec == 1 && model == 6)
             ? (k == 3 ? 7.67 : k == 4 ? 138.3 : (k == 6 || k == 7) ? 0.75 : 0.0)
             : 0.0;
}

// Rush-Larsen gate update g <- g_inf - (g_inf - g) * exp(-dt / tau) (e.g.
// tp06_2d.py:307-309) of a gate stored as d = g - g0.  FP64 path (g0 = 0):
// that expression.  FP32 path: d advances by the increment,
// d + ((g_inf - g0) - d) * (1 - e) with 1 - e = -expm1(-dt / tau).  A float
// gate near its steady value stops moving once the increment falls below half
// an ulp of the gate: TP06's j at rest, dt / tau_j = 1.4e-4, stalled anywhere
// within 2.1e-4 of j_inf, its availability drifted by 2.7e-4 after 780 ms and
// the conduction velocity with it.  Stored as d the stall band is half an ulp
// of d over dt / tau (6.5e-6 for |d| < 0.031), and 1 - e taken from expm1
// keeps the digits a float e would lose for slow gates.
template <typename R>
__device__ __forceinline__ R rl_gate(R d, R inf, R tau, R dt, R g0) {
  if constexpr (sizeof(R) == 4) {
    const R em = -expm1f(-dt / tau);
    return d + ((inf - g0) - d) * em;
  } else {
    return inf - (inf - d) * fexp(dv(-dt, tau));
  }
}

// ---- Aliev-Panfilov  aliev_panfilov_2d.py:133-203 ---------------------------
template <typename R, typename SP>
__device__ __forceinline__ R ionic_ap(R u, R un, SP& S, int64_t c,
                                      R dt, const Params& P) {
  const R a = par<R>(P, 0), k = par<R>(P, 1), eap = par<R>(P, 2), mu_1 = par<R>(P, 3),
          mu_2 = par<R>(P, 4);
  auto&& vref = ref<R>(S, 0, c);
  R v = vref;
  // calc_v: v += (-dt*(eap + (mu_1*v)/(mu_2 + u)) * (v + k*u*(u - a - 1.)))
  v += (-dt * (eap + (mu_1 * v) / (mu_2 + u)) * (v + k * u * (u - a - R(1))));
  vref = v;
  // u_new += dt*(-k*u*(u - a)*(u - 1.) - u*v)
  return un + dt * (-k * u * (u - a) * (u - R(1)) - u * v);
}

// ---- Barkley  barkley_2d.py:113-166 ----------------------------------------
template <typename R, typename SP>
__device__ __forceinline__ R ionic_barkley(R u, R un, SP& S,
                                           int64_t c, R dt, const Params& P) {
  const R a = par<R>(P, 0), b = par<R>(P, 1), eap = par<R>(P, 2);
  auto&& vref = ref<R>(S, 0, c);
  R v = vref;
  v += dt * (u - v);
  vref = v;
  return un + dt * (u * (R(1) - u) * (u - (v + b) / a)) / eap;
}

// ---- Mitchell-Schaeffer  mitchell_schaeffer_2d.py:102-217 ------------------
template <typename R, typename SP>
__device__ __forceinline__ R ionic_ms(R u, R un, SP& S, int64_t c,
                                      R dt, const Params& P) {
  const R tau_close = par<R>(P, 0), tau_open = par<R>(P, 1), tau_in = par<R>(P, 2),
          tau_out = par<R>(P, 3), u_gate = par<R>(P, 4);
  auto&& href = ref<R>(S, 0, c);
  R h = href;
  if constexpr (sizeof(R) == 4) {
    // float path: multiplications by 1 / tau_open and 1 / tau_close, rounded
    // once on the host (parameters 5 and 6), instead of a reciprocal per
    // node of the tau the select picks
    const R inv_open = par<R>(P, 5), inv_close = par<R>(P, 6);
    h += (u < u_gate) ? dt * (R(1) - h) * inv_open : dt * (-h) * inv_close;
  } else {
    h += (u < u_gate) ? dt * (R(1) - h) / tau_open : dt * (-h) / tau_close;
  }
  href = h;
  R C = (u * u) * (R(1) - u);  // (u**2)*(1-u), u**2 -> u*u
  R J_in = h * C / tau_in;
  R J_out = -u / tau_out;
  return un + dt * (J_in + J_out);
}

// ---- Fenton-Karma  fenton_karma_2d.py:143-328 ------------------------------
template <typename R, typename SP>
__device__ __forceinline__ R ionic_fk(R u, R un, SP& S, int64_t c,
                                      R dt, const Params& P) {
  const R tau_d = par<R>(P, 0), tau_o = par<R>(P, 1), tau_r = par<R>(P, 2),
          tau_si = par<R>(P, 3), tau_v_m = par<R>(P, 4), tau_v_p = par<R>(P, 5),
          tau_w_m = par<R>(P, 6), tau_w_p = par<R>(P, 7), k = par<R>(P, 8),
          u_c = par<R>(P, 9), uc_si = par<R>(P, 10);
  auto&& vref = ref<R>(S, 0, c);
  auto&& wref = ref<R>(S, 1, c);
  R v = vref, w = wref;
  const R H1 = (u_c - u) >= 0 ? R(1) : R(0);
  const R H2 = (u - u_c) >= 0 ? R(1) : R(0);
  v += dt * (H1 * (R(1) - v) / tau_v_m - H2 * v / tau_v_p);
  w += dt * (H1 * (R(1) - w) / tau_w_m - H2 * w / tau_w_p);
  vref = v;
  wref = w;
  const R J_fi = -(v * H2 * (R(1) - u) * (u - u_c)) / tau_d;
  const R J_so = u * H1 / tau_o + H2 / tau_r;
  // calc_Jsi receives v (fenton_karma_2d.py:326)
  const R J_si = -v * (R(1) + ftanh(k * (u - uc_si))) / (R(2) * tau_si);
  return un + dt * (-J_fi - J_so - J_si);
}

// ---- Bueno-Orovio  bueno_orovio_2d.py:183-476 -------------------------------
template <typename R, typename SP>
__device__ __forceinline__ R ionic_bo(R u, R un, SP& S, int64_t c,
                                      R dt, const Params& P) {
  const R u_o = par<R>(P, 0), u_u = par<R>(P, 1), theta_v = par<R>(P, 2),
          theta_w = par<R>(P, 3), theta_v_m = par<R>(P, 4), theta_o = par<R>(P, 5),
          tau_v1_m = par<R>(P, 6), tau_v2_m = par<R>(P, 7), tau_v_p = par<R>(P, 8),
          tau_w1_m = par<R>(P, 9), tau_w2_m = par<R>(P, 10), k_w_m = par<R>(P, 11),
          u_w_m = par<R>(P, 12), tau_w_p = par<R>(P, 13), tau_fi = par<R>(P, 14),
          tau_o1 = par<R>(P, 15), tau_o2 = par<R>(P, 16), tau_so1 = par<R>(P, 17),
          tau_so2 = par<R>(P, 18), k_so = par<R>(P, 19), u_so = par<R>(P, 20),
          tau_s1 = par<R>(P, 21), tau_s2 = par<R>(P, 22), k_s = par<R>(P, 23),
          u_s = par<R>(P, 24), tau_si = par<R>(P, 25), tau_w_inf = par<R>(P, 26),
          w_inf_ = par<R>(P, 27);
  auto&& vref = ref<R>(S, 0, c);
  auto&& wref = ref<R>(S, 1, c);
  auto&& sref = ref<R>(S, 2, c);
  R v = vref, w = wref, s = sref;
  // calc_v(v, u, dt, theta_v, calc_v_inf(u), calc_tau_v_m(u), tau_v_p)
  {
    const R v_inf = u < theta_v_m ? R(1) : R(0);
    const R tau_v_m = (u - theta_v_m) < 0 ? tau_v1_m : tau_v2_m;
    const R v_ = (u - theta_v) < 0 ? (v_inf - v) / tau_v_m : -v / tau_v_p;
    v += dt * v_;
  }
  // calc_w(w, u, dt, theta_w, calc_w_inf(u), calc_tau_w_m(u), tau_w_p)
  {
    const R w_inf = (u - theta_o) < 0 ? R(1) - u / tau_w_inf : w_inf_;
    const R tau_w_m = tau_w1_m + (tau_w2_m - tau_w1_m) *
                                     (R(1) + ftanh(k_w_m * (u - u_w_m))) / R(2);
    const R w_ = (u - theta_w) < 0 ? (w_inf - w) / tau_w_m : -w / tau_w_p;
    w += dt * w_;
  }
  // calc_s(s, u, dt, calc_tau_s(u), k_s, u_s)
  {
    const R tau_s = (u - theta_w) < 0 ? tau_s1 : tau_s2;
    s += dt * ((R(1) + ftanh(k_s * (u - u_s))) / R(2) - s) / tau_s;
  }
  vref = v;
  wref = w;
  sref = s;
  const R Hv = (u - theta_v) >= 0 ? R(1) : R(0);
  const R J_fi = -v * Hv * (u - theta_v) * (u_u - u) / tau_fi;
  const R tau_o = (u - theta_o) < 0 ? tau_o1 : tau_o2;
  const R tau_so =
      tau_so1 + (tau_so2 - tau_so1) * (R(1) + ftanh(k_so * (u - u_so))) / R(2);
  const R Hw = (u - theta_w) >= 0 ? R(1) : R(0);
  const R J_so = (u - u_o) * (R(1) - Hw) / tau_o + Hw / tau_so;
  const R J_si = -Hw * w * s / tau_si;
  return un + dt * (-J_fi - J_so - J_si);
}

// ---- Luo-Rudy 1991  luo_rudy91_2d.py:158-510 --------------------------------
template <typename R>
__device__ __forceinline__ R lr91_gate(R var, R dt, R alpha, R beta) {
  const R tau = R(1) / (alpha + beta);
  const R inf = alpha / (alpha + beta);
  var += dt * (inf - var) / tau;
  return var;
}

template <typename R, typename SP, typename U = R>
__device__ __forceinline__ U ionic_lr91(U u_in, U un, SP& S, int64_t c, R dt,
                                        const Params& P) {
  const R u = (R)u_in;  // the potential in the arithmetic type
  const R gna = par<R>(P, 0), gsi = par<R>(P, 1), gk = par<R>(P, 2), gk1 = par<R>(P, 3),
          gkp = par<R>(P, 4), gb = par<R>(P, 5), ko = par<R>(P, 6), ki = par<R>(P, 7),
          nai = par<R>(P, 8), nao = par<R>(P, 9), R_ = par<R>(P, 11), T = par<R>(P, 12),
          F = par<R>(P, 13), PR_NaK = par<R>(P, 14);
  // E_Na = R*T/F*np.log(nao/nai) (computed once per launch in the reference)
  const R E_Na = R_ * T / F * flog(nao / nai);
  auto&& mr = ref<R>(S, 0, c);
  auto&& hr = ref<R>(S, 1, c);
  auto&& jr = ref<R>(S, 2, c);
  auto&& dr = ref<R>(S, 3, c);
  auto&& fr = ref<R>(S, 4, c);
  auto&& xr = ref<R>(S, 5, c);
  auto&& cair = ref<R>(S, 6, c);
  // calc_ina
  R ina;
  {
    R alpha_h = 0, beta_h = 0, beta_J = 0, alpha_J = 0;
    if (u >= R(-40.0)) {
      beta_h = R(1.0) / (R(0.13) * (R(1) + fexp((u + R(10.66)) / R(-11.1))));
      beta_J = R(0.3) * fexp(R(-2.535 * 1e-07) * u) /
               (R(1) + fexp(R(-0.1) * (u + R(32))));
    } else {
      alpha_h = R(0.135) * fexp((R(80) + u) / R(-6.8));
      beta_h = R(3.56) * fexp(R(0.079) * u) +
               R(3.1 * 100000.0) * fexp(R(0.35) * u);
      beta_J = R(0.1212) * fexp(R(-0.01052) * u) /
               (R(1) + fexp(R(-0.1378) * (u + R(40.14))));
      alpha_J = (R(-1.2714 * 100000.0) * fexp(R(0.2444) * u) -
                 R(3.474 * 1e-05) * fexp(R(-0.04391) * u)) *
                (u + R(37.78)) / (R(1) + fexp(R(0.311) * (u + R(79.23))));
    }
    R alpha_m;
    if constexpr (sizeof(R) == 4)
      alpha_m = R(3.2) / exprel(R(-0.1) * (u + R(47.13)));
    else
      alpha_m = R(0.32) * (u + R(47.13)) / (R(1) - fexp(R(-0.1) * (u + R(47.13))));
    const R beta_m = R(0.08) * fexp(-u / R(11));
    const R m = lr91_gate<R>(mr, dt, alpha_m, beta_m);
    const R h = lr91_gate<R>(hr, dt, alpha_h, beta_h);
    const R j = lr91_gate<R>(jr, dt, alpha_J, beta_J);
    mr = m;
    hr = h;
    jr = j;
    ina = gna * m * m * m * h * j * (u - E_Na);
  }
  // calc_isk: the current uses the old d, f
  R isi;
  {
    R d = dr, f = fr, cai = cair;
    const R E_Si = R(7.7) - R(13.0287) * flog(cai);
    isi = gsi * d * f * (u - E_Si);
    const R alpha_d = R(0.095) * fexp(R(-0.01) * (u - R(5))) /
                      (R(1) + fexp(R(-0.072) * (u - R(5))));
    const R beta_d = R(0.07) * fexp(R(-0.017) * (u + R(44))) /
                     (R(1) + fexp(R(0.05) * (u + R(44))));
    const R alpha_f = R(0.012) * fexp(R(-0.008) * (u + R(28))) /
                      (R(1) + fexp(R(0.15) * (u + R(28))));
    const R beta_f = R(0.0065) * fexp(R(-0.02) * (u + R(30))) /
                     (R(1) + fexp(R(-0.2) * (u + R(30))));
    d = lr91_gate<R>(d, dt, alpha_d, beta_d);
    f = lr91_gate<R>(f, dt, alpha_f, beta_f);
    cai += dt * (R(-0.0001) * isi + R(0.07) * (R(0.0001) - cai));
    dr = d;
    fr = f;
    cair = cai;
  }
  // calc_ik: the current uses the old x
  R ik;
  {
    const R E_K = R_ * T / F * flog((ko + PR_NaK * nao) / (ki + PR_NaK * nai));
    const R G_K = gk * fsqrt(ko / R(5.4));
    R Xi;
    if (u > R(-100)) {
      if constexpr (sizeof(R) == 4)
        Xi = R(2.837 * 0.04) * exprel(R(0.04) * (u + R(77))) / fexp(R(0.04) * (u + R(35)));
      else
        Xi = R(2.837) * (fexp(R(0.04) * (u + R(77))) - R(1)) /
             ((u + R(77)) * fexp(R(0.04) * (u + R(35))));
    } else {
      Xi = R(1);
    }
    const R x = xr;
    ik = G_K * x * Xi * (u - E_K);
    const R alpha_x = R(0.0005) * fexp(R(0.083) * (u + R(50))) /
                      (R(1) + fexp(R(0.057) * (u + R(50))));
    const R beta_x = R(0.0013) * fexp(R(-0.06) * (u + R(20))) /
                     (R(1) + fexp(R(-0.04) * (u + R(20))));
    xr = lr91_gate<R>(x, dt, alpha_x, beta_x);
  }
  const R E_K1 = R_ * T / F * flog(ko / ki);
  // calc_ik1
  const R alpha_K1 = R(1.02) / (R(1) + fexp(R(0.2385) * (u - E_K1 - R(59.215))));
  const R beta_K1 = (R(0.49124) * fexp(R(0.08032) * (u - E_K1 + R(5.476))) +
                     fexp(R(0.06175) * (u - E_K1 - R(594.31)))) /
                    (R(1) + fexp(R(-0.5143) * (u - E_K1 + R(4.753))));
  const R K_1x = alpha_K1 / (alpha_K1 + beta_K1);
  const R G_K1 = gk1 * fsqrt(ko / R(5.4));
  const R ik1 = G_K1 * K_1x * (u - E_K1);
  // calc_ikp, calc_ib
  const R K_p = R(1.0) / (R(1) + fexp((R(7.488) - u) / R(5.98)));
  const R ikp = gkp * K_p * (u - E_K1);
  const R ib = gb * (u + R(59.87));
  const R ik1t = ik1 + ikp + ib;
  return un - dt * (ina + isi + ik1t + ik);
}

// ---- TP06  tp06_2d.py:242-1103 -----------------------------------------------
// state order (tp06_2d.py:1073 call): cai casr cass nai Ki m h j xr1 xr2 xs r
// s d f f2 fcass rr oo ; the first five are double in both precisions.
template <typename R, typename SP, typename U = R>
__device__ __forceinline__ U ionic_tp06(U u_in, U un, SP& S, int64_t c, R dt,
                                        const Params& P) {
  const R u = (R)u_in;  // the potential in the arithmetic type
  typedef double D;
  const R ko = par<R>(P, 0), cao = par<R>(P, 1), nao = par<R>(P, 2);
  const D Vc = P.v[3], Vsr = P.v[4], Vss = P.v[5];
  const D Bufsr = P.v[8], Kbufsr = P.v[9], Bufss = P.v[10], Kbufss = P.v[11],
          Vmaxup = P.v[12], Kup = P.v[13], Vrel = P.v[14], k1_ = P.v[15],
          k2_ = P.v[16], k3 = P.v[17], k4 = P.v[18], EC = P.v[19],
          maxsr = P.v[20], minsr = P.v[21], Vleak = P.v[22], Vxfer = P.v[23];
  const R R_ = par<R>(P, 24), F = par<R>(P, 25), T = par<R>(P, 26), RTONF = par<R>(P, 27);
  const D CAPACITANCE = P.v[28], Fd = P.v[25];
  const R gkr = par<R>(P, 29), pKNa = par<R>(P, 30), gk1 = par<R>(P, 31),
          gna = par<R>(P, 32), gbna = par<R>(P, 33), KmK = par<R>(P, 34),
          KmNa = par<R>(P, 35), knak = par<R>(P, 36), gcal = par<R>(P, 37),
          gbca = par<R>(P, 38), knaca = par<R>(P, 39), KmNai = par<R>(P, 40),
          KmCa = par<R>(P, 41), ksat = par<R>(P, 42), n_ = par<R>(P, 43),
          gpca = par<R>(P, 44), KpCa = par<R>(P, 45), gpk = par<R>(P, 46),
          gto = par<R>(P, 47), gks = par<R>(P, 48);
  constexpr bool F64 = sizeof(R) == 8;
  const double* Dv = P.v + kTp06Derived;  // uniform quantities (complete_params)
  const D inverseVcF = Dv[kInvVcF];
  const D inversevssF2 = Dv[kInvVssF2];
  // x / (R T) for a node-dependent x
  auto overRT = [&](R x) -> R {
    if constexpr (F64) return div_rn(x, par<R>(P, kTp06Derived + kRT), par<R>(P, kTp06Derived + kInvRT));
    else return x / (R_ * T);
  };

  // nai, Ki: stored as differences from state_offset on the FP32 path
  typedef std::conditional_t<F64, double, float> Sc;
  constexpr D kNai0 = state_offset(6, F64 ? 0 : 1, 3), kKi0 = state_offset(6, F64 ? 0 : 1, 4);
  auto&& cair = ref<D>(S, 0, c);
  auto&& casrr = ref<D>(S, 1, c);
  auto&& cassr = ref<D>(S, 2, c);
  auto&& nair = ref<Sc>(S, 3, c);
  auto&& Kir = ref<Sc>(S, 4, c);
  // read where they are used: the balances at the end read them again
  // rather than keep them live across the currents
  const R cai = (R)(D)cair, cass = (R)(D)cassr;
  const R nai = (R)(F64 ? (D)(Sc)nair : (D)(Sc)nair + kNai0);
  const R Ki = (R)(F64 ? (D)(Sc)Kir : (D)(Sc)Kir + kKi0);

  const R Ek = RTONF * flog(dv(ko, Ki));
  const R Ena = RTONF * flog(dv(nao, nai));
  const R Eks = RTONF * flog(dv(ko + pKNa * nao, Ki + pKNa * nai));
  const R Eca = R(0.5) * RTONF * flog(dv(cao, cai));

  // calc_ina
  R ina;
  {
    // h, j: stored as differences from state_offset on the FP32 path
    constexpr R kH0 = (R)state_offset(6, F64 ? 0 : 1, 6), kJ0 = (R)state_offset(6, F64 ? 0 : 1, 7);
    auto&& mr = ref<R>(S, 5, c);
    auto&& hr = ref<R>(S, 6, c);
    auto&& jr = ref<R>(S, 7, c);
    const R alpha_m = dv(R(1.0), R(1.0) + fexp(divc(R(-60.0) - u, R(5.0))));
    const R beta_m = dv(R(0.1), R(1.0) + fexp(divc(u + R(35.0), R(5.0)))) +
                     dv(R(0.1), R(1.0) + fexp(divc(u - R(50.0), R(200.0))));
    const R tau_m = alpha_m * beta_m;
    const R m_inf = dv(R(1.0), (R(1.0) + fexp(divc(R(-56.86) - u, R(9.03)))) *
                              (R(1.0) + fexp(divc(R(-56.86) - u, R(9.03)))));
    R alpha_h, beta_h;
    if (u >= R(-40.0)) {
      alpha_h = R(0.0);
      beta_h = dv(R(0.77), R(0.13) * (R(1.0) + fexp(divc(-(u + R(10.66)), R(11.1)))));
    } else {
      alpha_h = R(0.057) * fexp(divc(-(u + R(80.0)), R(6.8)));
      beta_h = R(2.7) * fexp(R(0.079) * u) + R(310000.0) * fexp(R(0.3485) * u);
    }
    const R tau_h = dv(R(1.0), alpha_h + beta_h);
    const R h_inf = dv(R(1.0), (R(1.0) + fexp(divc(u + R(71.55), R(7.43)))) *
                              (R(1.0) + fexp(divc(u + R(71.55), R(7.43)))));
    R alpha_j, beta_j;
    if (u >= R(-40.0)) {
      alpha_j = R(0.0);
      beta_j = dv(R(0.6) * fexp(R(0.057) * u),
                R(1.0) + fexp(R(-0.1) * (u + R(32.0))));
    } else {
      alpha_j = dv((R(-25428.0) * fexp(R(0.2444) * u) -
                    R(6.948e-06) * fexp(R(-0.04391) * u)) *
                       (u + R(37.78)),
                   R(1.0) + fexp(R(0.311) * (u + R(79.23))));
      beta_j = dv(R(0.02424) * fexp(R(-0.01052) * u),
                R(1.0) + fexp(R(-0.1378) * (u + R(40.14))));
    }
    const R tau_j = dv(R(1.0), alpha_j + beta_j);
    const R j_inf = h_inf;
    const R m = m_inf - (m_inf - mr) * fexp(dv(-dt, tau_m));
    const R hd = rl_gate<R>(hr, h_inf, tau_h, dt, kH0);
    const R jd = rl_gate<R>(jr, j_inf, tau_j, dt, kJ0);
    mr = m;
    hr = hd;
    jr = jd;
    const R h = F64 ? hd : hd + kH0, j = F64 ? jd : jd + kJ0;
    ina = gna * m * m * m * h * j * (u - Ena);
  }
  // calc_ical
  R ical;
  {
    auto&& dr = ref<R>(S, 13, c);
    auto&& fr = ref<R>(S, 14, c);
    auto&& f2r = ref<R>(S, 15, c);
    auto&& fcr = ref<R>(S, 16, c);
    const R d_inf = dv(R(1.0), R(1.0) + fexp(divc(R(-8) - u, R(7.5))));
    const R Ad = dv(R(1.4), R(1.0) + fexp(divc(R(-35) - u, R(13)))) + R(0.25);
    const R Bd = dv(R(1.4), R(1.0) + fexp(divc(u + R(5), R(5))));
    const R Cd = dv(R(1.0), R(1.0) + fexp(divc(R(50) - u, R(20))));
    const R tau_d = Ad * Bd + Cd;
    const R f_inf = dv(R(1.0), R(1.0) + fexp(divc(u + R(20), R(7))));
    const R Af = R(1102.5) * fexp(divc(-(u + R(27)) * (u + R(27)), R(225)));
    const R Bf = dv(R(200.0), R(1) + fexp(divc(R(13) - u, R(10.0))));
    const R Cf = dv(R(180.0), R(1) + fexp(divc(u + R(30), R(10)))) + R(20);
    const R tau_f = Af + Bf + Cf;
    const R f2_inf = dv(R(0.67), R(1.0) + fexp(divc(u + R(35), R(7)))) + R(0.33);
    const R Af2 = R(600) * fexp(divc(-(u + R(25)) * (u + R(25)), R(170)));
    const R Bf2 = dv(R(31), R(1.0) + fexp(divc(R(25) - u, R(10))));
    const R Cf2 = dv(R(16), R(1.0) + fexp(divc(u + R(30), R(10))));
    const R tau_f2 = Af2 + Bf2 + Cf2;
    const R fcass_inf = dv(R(0.6), R(1) + divc(cass, R(0.05)) * divc(cass, R(0.05))) + R(0.4);
    const R tau_fcass = dv(R(80.0), R(1) + divc(cass, R(0.05)) * divc(cass, R(0.05))) + R(2.0);
    const R d = d_inf - (d_inf - dr) * fexp(dv(-dt, tau_d));
    const R f = f_inf - (f_inf - fr) * fexp(dv(-dt, tau_f));
    const R f2 = f2_inf - (f2_inf - f2r) * fexp(dv(-dt, tau_f2));
    const R fcass = fcass_inf - (fcass_inf - fcr) * fexp(dv(-dt, tau_fcass));
    dr = d;
    fr = f;
    f2r = f2;
    fcr = fcass;
    if constexpr (sizeof(R) == 4) {
      // (u - 15) / (e^X - 1) = 1 / ((2F/RT) * exprel(X)),  X = 2(u - 15)F/RT
      const R X = R(2) * (u - R(15)) * F / (R_ * T);
      ical = gcal * d * f * f2 * fcass * R(4) * par<R>(P, kTp06Derived + kFFoverRT) *
             (R(0.25) * fexp(X) * cass - cao) / (R(2) * F / (R_ * T) * exprel(X));
    } else {
      const R e = fexp(overRT(R(2) * (u - R(15)) * F));
      ical = dv(gcal * d * f * f2 * fcass * R(4) * (u - R(15)) *
                    par<R>(P, kTp06Derived + kFFoverRT) * (R(0.25) * e * cass - cao),
                e - R(1.0));
    }
  }
  // calc_ito
  R ito;
  {
    auto&& rr_ = ref<R>(S, 11, c);
    auto&& sr = ref<R>(S, 12, c);
    const R r_inf = dv(R(1.0), R(1.0) + fexp(divc(R(20) - u, R(6.0))));
    const R s_inf = dv(R(1.0), R(1.0) + fexp(divc(u + R(20), R(5.0))));
    const R tau_r = R(9.5) * fexp(divc(-(u + R(40.0)) * (u + R(40.0)), R(1800.0))) + R(0.8);
    const R tau_s = R(85.0) * fexp(divc(-(u + R(45.0)) * (u + R(45.0)), R(320.0))) +
                    dv(R(5.0), R(1.0) + fexp(divc(u - R(20.0), R(5.0)))) + R(3.0);
    const R s = s_inf - (s_inf - sr) * fexp(dv(-dt, tau_s));
    const R r = r_inf - (r_inf - rr_) * fexp(dv(-dt, tau_r));
    sr = s;
    rr_ = r;
    ito = gto * r * s * (u - Ek);
  }
  // calc_ikr
  R ikr;
  {
    auto&& x1r = ref<R>(S, 8, c);
    auto&& x2r = ref<R>(S, 9, c);
    const R xr1_inf = dv(R(1.0), R(1.0) + fexp(divc(R(-26.0) - u, R(7.0))));
    const R axr1 = dv(R(450.0), R(1.0) + fexp(divc(R(-45.0) - u, R(10.0))));
    const R bxr1 = dv(R(6.0), R(1.0) + fexp(divc(u - R(-30.0), R(11.5))));
    const R tau_xr1 = axr1 * bxr1;
    const R xr2_inf = dv(R(1.0), R(1.0) + fexp(divc(u - R(-88.0), R(24.0))));
    const R axr2 = dv(R(3.0), R(1.0) + fexp(divc(R(-60.0) - u, R(20.0))));
    const R bxr2 = dv(R(1.12), R(1.0) + fexp(divc(u - R(60.0), R(20.0))));
    const R tau_xr2 = axr2 * bxr2;
    const R xr1 = xr1_inf - (xr1_inf - x1r) * fexp(dv(-dt, tau_xr1));
    const R xr2 = xr2_inf - (xr2_inf - x2r) * fexp(dv(-dt, tau_xr2));
    x1r = xr1;
    x2r = xr2;
    ikr = gkr * par<R>(P, kTp06Derived + kSqrtKo54) * xr1 * xr2 * (u - Ek);
  }
  // calc_iks
  R iks;
  {
    auto&& xsr = ref<R>(S, 10, c);
    const R Axs = dv(R(1400.0), fsqrt(R(1.0) + fexp(divc(R(5.0) - u, R(6)))));
    const R Bxs = dv(R(1.0), R(1.0) + fexp(divc(u - R(35.0), R(15.0))));
    const R tau_xs = Axs * Bxs + R(80);
    const R xs_inf = dv(R(1.0), R(1.0) + fexp(divc(R(-5.0) - u, R(14.0))));
    const R xs = xs_inf - (xs_inf - xsr) * fexp(dv(-dt, tau_xs));
    xsr = xs;
    iks = gks * xs * xs * (u - Eks);
  }
  // calc_ik1
  const R ak1 = dv(R(0.1), R(1.0) + fexp(R(0.06) * (u - Ek - R(200))));
  const R bk1 = dv(R(3.0) * fexp(R(0.0002) * (u - Ek + R(100))) +
                      fexp(R(0.1) * (u - Ek - R(10))),
                  R(1.0) + fexp(R(-0.5) * (u - Ek)));
  const R rec_iK1 = dv(ak1, ak1 + bk1);
  const R ik1 = gk1 * rec_iK1 * (u - Ek);
  // calc_inaca
  const R e_n1 = fexp(overRT((n_ - R(1)) * u * F));
  const R inaca =
      knaca *
      par<R>(P, kTp06Derived + kInvNaCaK) * par<R>(P, kTp06Derived + kInvKmCaCao) * dv(R(1.0), R(1) + ksat * e_n1) *
      (fexp(overRT(n_ * u * F)) * nai * nai * nai * cao - e_n1 * nao * nao * nao * cai * R(2.5));
  // calc_inak
  const R rec_iNaK = dv(R(1.0), R(1.0) + R(0.1245) * fexp(overRT(R(-0.1) * u * F)) +
                                     R(0.0353) * fexp(overRT(-u * F)));
  const R inak =
      knak * par<R>(P, kTp06Derived + kKoOverKoKmK) * dv(nai, nai + KmNa) * rec_iNaK;
  const R ipca = dv(gpca * cai, KpCa + cai);
  const R rec_ipK = dv(R(1.0), R(1.0) + fexp(divc(R(25) - u, R(5.98))));
  const R ipk = gpk * rec_ipK * (u - Ek);
  const R ibna = gbna * (u - Ena);
  const R ibca = gbca * (u - Eca);
  // calc_irel (double: feeds the SR/SS calcium balance)
  D irel;
  {
    auto&& rrr = ref<R>(S, 17, c);
    auto&& oor = ref<R>(S, 18, c);
    const D casrD = casrr, cassD = cassr;
    const D kCaSR = maxsr - dv(maxsr - minsr, 1 + dv(EC, casrD) * dv(EC, casrD));
    const D k1 = dv(k1_, kCaSR);
    const D k2 = k2_ * kCaSR;
    D rr = (D)rrr;
    const D drr = k4 * (1 - rr) - k2 * cassD * rr;
    rr += (D)dt * drr;
    const D oo = dv(k1 * cassD * cassD * rr, k3 + k1 * cassD * cassD);
    rrr = (R)rr;
    oor = (R)oo;
    irel = Vrel * oo * (casrD - cassD);
  }
  const D caiD = cair, casrD = casrr, cassD = cassr;
  const D ileak = Vleak * (casrD - caiD);
  const D iup = dv(Vmaxup, 1.0 + dv(Kup * Kup, caiD * caiD));
  const D ixfer = Vxfer * (cassD - caiD);
  const D dtD = (D)dt;
  // calc_casr: closed-form buffer quadratic
  {
    const D CaCSQN = dv(Bufsr * casrD, casrD + Kbufsr);
    const D dCaSR = dtD * (iup - irel - ileak);
    const D bjsr = Bufsr - CaCSQN - dCaSR - casrD + Kbufsr;
    const D cjsr = Kbufsr * (CaCSQN + dCaSR + casrD);
    casrr = (sqrt(bjsr * bjsr + 4 * cjsr) - bjsr) / 2;
  }
  // calc_cass
  {
    const D CaSSBuf = dv(Bufss * cassD, cassD + Kbufss);
    const D dCaSS = dtD * (-ixfer * Dv[kVcOverVss] + irel * Dv[kVsrOverVss] +
                           -(D)ical * inversevssF2 * CAPACITANCE);
    const D bcss = Bufss - CaSSBuf - dCaSS - cassD + Kbufss;
    const D ccss = Kbufss * (CaSSBuf + dCaSS + cassD);
    cassr = (sqrt(bcss * bcss + 4 * ccss) - bcss) / 2;
  }
  // calc_cai's result is discarded by the tuple store (tp06_2d.py:1098).
  // calc_nai / calc_ki
  {
    const D dNai = -((D)ina + (D)ibna + 3 * (D)inak + 3 * (D)inaca) * inverseVcF *
                   CAPACITANCE;
    nair = (Sc)((D)(Sc)nair + dtD * dNai);
    const D dKi = -((D)ik1 + (D)ito + (D)ikr + (D)iks - 2 * (D)inak + (D)ipk) *
                  inverseVcF * CAPACITANCE;
    Kir = (Sc)((D)(Sc)Kir + dtD * dKi);
  }
  return un - dt * (ikr + iks + ik1 + ito + ina + ibna + ical + ibca + inak +
                    inaca + ipca + ipk);
}

// ---- Courtemanche  courtemanche_2d.py:180-641 -------------------------------
// state order = the kernel's parameter order (courtemanche_2d.py:568-569):
// nai ki cai caup carel m h j_ d f oa oi ua ui xs xr fca irel vrel urel wrel;
// the model passes its xr array in the xs slot and vice versa.
template <typename R>
__device__ __forceinline__ R crn_gate(R x, R x_inf, R tau_x, R dt) {
  return x_inf - (x_inf - x) * fexp(-dt / tau_x);
}

template <typename R, typename SP, typename U = R>
__device__ __forceinline__ U ionic_crn(U u_in, U un, SP& S, int64_t c, R dt,
                                        const Params& P) {
  const R u = (R)u_in;  // the potential in the arithmetic type
  const R gna = par<R>(P, 0), gnab = par<R>(P, 1), gk1 = par<R>(P, 2), gks = par<R>(P, 4),
          gto = par<R>(P, 5), gcal = par<R>(P, 6), gcab = par<R>(P, 7),
          gkur_coeff = par<R>(P, 8), F = par<R>(P, 9), T = par<R>(P, 10),
          R_ = par<R>(P, 11), Vj = par<R>(P, 13), Vup = par<R>(P, 14),
          Vrel = par<R>(P, 15), ibk = par<R>(P, 16), cao = par<R>(P, 17),
          nao = par<R>(P, 18), ko = par<R>(P, 19), caupmax = par<R>(P, 20),
          kup = par<R>(P, 21), kmnai = par<R>(P, 22), kmko = par<R>(P, 23),
          kmnancx = par<R>(P, 24), kmcancx = par<R>(P, 25), ksatncx = par<R>(P, 26),
          kmcmdn = par<R>(P, 27), kmtrpn = par<R>(P, 28), kmcsqn = par<R>(P, 29),
          trpnmax = par<R>(P, 30), cmdnmax = par<R>(P, 31), csqnmax = par<R>(P, 32),
          inacamax = par<R>(P, 33), inakmax = par<R>(P, 34), ipcamax = par<R>(P, 35),
          krel = par<R>(P, 36), iupmax = par<R>(P, 37), kq10 = par<R>(P, 38);
  auto&& nair = ref<R>(S, 0, c);
  auto&& kir = ref<R>(S, 1, c);
  auto&& cair = ref<R>(S, 2, c);
  auto&& caupr = ref<R>(S, 3, c);
  auto&& carelr = ref<R>(S, 4, c);
  const R nai = nair, ki = kir, cai = cair, caup = caupr, carel = carelr;
  // calc_equilibrum_potentials
  const R ena = R_ * T / F * flog(nao / nai);
  const R ek = R_ * T / F * flog(ko / ki);
  const R eca = R_ * T / (R(2) * F) * flog(cao / cai);
  // calc_gating_m / h / j
  R m, h, j;
  {
    auto&& mr = ref<R>(S, 5, c);
    R am;
    if constexpr (sizeof(R) == 4) {
      am = R(3.2) / exprel(R(-0.1) * (u + R(47.13)));
    } else {
      if (u == R(-47.13)) {
        am = R(3.2);
      } else {
        am = R(0.32) * (u + R(47.13)) / (R(1) - fexp(R(-0.1) * (u + R(47.13))));
      }
    }
    const R bm = R(0.08) * fexp(-u / R(11));
    const R m_inf = am / (am + bm);
    const R tau_m = R(1) / (am + bm);
    m = crn_gate<R>(mr, m_inf, tau_m, dt);
    mr = m;
  }
  {
    auto&& hr = ref<R>(S, 6, c);
    R ah, bh;
    if (u >= R(-40)) {
      ah = R(0);
      bh = R(1) / (R(0.13) * (R(1) + fexp(-(u + R(10.66)) / R(11.1))));
    } else {
      ah = R(0.135) * fexp(-(R(80) + u) / R(6.8));
      bh = R(3.56) * fexp(R(0.079) * u) + R(310000) * fexp(R(0.35) * u);
    }
    const R h_inf = ah / (ah + bh);
    const R tau_h = R(1) / (ah + bh);
    h = crn_gate<R>(hr, h_inf, tau_h, dt);
    hr = h;
  }
  {
    auto&& jr = ref<R>(S, 7, c);
    R aj, bj;
    if (u >= R(-40)) {
      aj = R(0);
      bj = R(0.3) * fexp(R(-2.535e-07) * u) / (R(1) + fexp(R(-0.1) * (u + R(32))));
    } else {
      aj = (R(-127140) * fexp(R(0.2444) * u) - R(3.474e-05) * fexp(R(-0.04391) * u)) *
           (u + R(37.78)) / (R(1) + fexp(R(0.311) * (u + R(79.23))));
      bj = R(0.1212) * fexp(R(-0.01052) * u) / (R(1) + fexp(R(-0.1378) * (u + R(40.14))));
    }
    const R j_inf = aj / (aj + bj);
    const R tau_j = R(1) / (aj + bj);
    j = j_inf - (j_inf - jr) * fexp(-dt / tau_j);
    jr = j;
  }
  const R ina = gna * (m * (m * m)) * h * j * (u - ena);
  const R ik1 = gk1 * (u - ek) / (R(1) + fexp(R(0.07) * (u + R(80))));
  // calc_ito
  R ito;
  {
    auto&& oar = ref<R>(S, 10, c);
    auto&& oir = ref<R>(S, 11, c);
    const R ao = R(0.65) / (fexp(-(u + R(10)) / R(8.5)) + fexp(-(u - R(30)) / R(59.0)));
    const R bo = R(0.65) / (R(2.5) + fexp((u + R(82)) / R(17.0)));
    const R tau_o = R(1) / (kq10 * (ao + bo));
    const R o_inf = R(1) / (R(1) + fexp(-(u + R(20.47)) / R(17.54)));
    const R aoi = R(1) / (R(18.53) + fexp((u + R(113.7)) / R(10.95)));
    const R boi = R(1) / (R(35.56) + fexp(-(u + R(1.26)) / R(7.44)));
    const R tau_oi = R(1) / (kq10 * (aoi + boi));
    const R oi_inf = R(1) / (R(1) + fexp((u + R(43.1)) / R(5.3)));
    const R oa = crn_gate<R>(oar, o_inf, tau_o, dt);
    const R oi = crn_gate<R>(oir, oi_inf, tau_oi, dt);
    oar = oa;
    oir = oi;
    ito = gto * (oa * (oa * oa)) * oi * (u - ek);
  }
  // calc_ikur
  R ikur;
  {
    auto&& uar = ref<R>(S, 12, c);
    auto&& uir = ref<R>(S, 13, c);
    const R gkur = R(0.005) + R(0.05) / (R(1) + fexp(-(u - R(15)) / R(13.0)));
    const R aua = R(0.65) / (fexp(-(u + R(10)) / R(8.5)) + fexp(-(u - R(30)) / R(59.0)));
    const R bua = R(0.65) / (R(2.5) + fexp((u + R(82)) / R(17.0)));
    const R tau_ua = R(1) / (kq10 * (aua + bua));
    const R ua_inf = R(1) / (R(1) + fexp(-(u + R(30.3)) / R(9.6)));
    const R aui = R(1) / (R(21) + fexp(-(u - R(185)) / R(28.0)));
    const R bui = fexp((u - R(158)) / R(16.0));
    const R tau_ui = R(1) / (kq10 * (aui + bui));
    const R ui_inf = R(1) / (R(1) + fexp((u - R(99.45)) / R(27.48)));
    const R ua = crn_gate<R>(uar, ua_inf, tau_ua, dt);
    const R ui = crn_gate<R>(uir, ui_inf, tau_ui, dt);
    uar = ua;
    uir = ui;
    ikur = gkur_coeff * gkur * (ua * (ua * ua)) * ui * (u - ek);
  }
  // calc_ikr on the kernel's xr slot (index 15), gkr hard-coded to 0.0294
  R ikr;
  {
    auto&& xrr = ref<R>(S, 15, c);
    const R gkr = R(0.0294);
    R axr, bxr;
    if constexpr (sizeof(R) == 4) {
      axr = R(0.0003 * 5) / exprel(-(u + R(14.1)) / R(5));
      bxr = R(7.3898e-05 * 5.1237) / exprel((u - R(3.3328)) / R(5.1237));
    } else {
      axr = R(0.0003) * (u + R(14.1)) / (R(1) - fexp(-(u + R(14.1)) / R(5)));
      bxr = R(7.3898e-05) * (u - R(3.3328)) / (fexp((u - R(3.3328)) / R(5.1237)) - R(1));
    }
    const R tau_xr = R(1) / (axr + bxr);
    const R xr_inf = R(1) / (R(1) + fexp(-(u + R(14.1)) / R(6.5)));
    const R xr = crn_gate<R>(xrr, xr_inf, tau_xr, dt);
    xrr = xr;
    ikr = gkr * xr * (u - ek) / (R(1) + fexp((u + R(15)) / R(22.4)));
  }
  // calc_iks on the kernel's xs slot (index 14)
  R iks;
  {
    auto&& xsr = ref<R>(S, 14, c);
    R axs, bxs;
    if constexpr (sizeof(R) == 4) {
      axs = R(4e-05 * 17) / exprel(-(u - R(19.9)) / R(17));
      bxs = R(3.5e-05 * 9) / exprel((u - R(19.9)) / R(9));
    } else {
      axs = R(4e-05) * (u - R(19.9)) / (R(1) - fexp(-(u - R(19.9)) / R(17)));
      bxs = R(3.5e-05) * (u - R(19.9)) / (fexp((u - R(19.9)) / R(9)) - R(1));
    }
    const R tau_xs = R(1) / (R(2) * (axs + bxs));
    const R xs_inf = R(1) / fsqrt(R(1) + fexp(-(u - R(19.9)) / R(12.7)));
    const R xs = crn_gate<R>(xsr, xs_inf, tau_xs, dt);
    xsr = xs;
    iks = gks * (xs * xs) * (u - ek);
  }
  // calc_ical
  R ical;
  {
    auto&& dr = ref<R>(S, 8, c);
    auto&& fr = ref<R>(S, 9, c);
    auto&& fcar = ref<R>(S, 16, c);
    R tau_d;
    if constexpr (sizeof(R) == 4) {
      const R y = -(u + R(10)) / R(6.24);
      tau_d = exprel(y) / (R(6.24 * 0.035) * (R(1) + fexp(y)));
    } else {
      tau_d = (R(1) - fexp(-(u + R(10)) / R(6.24))) /
              (R(0.035) * (u + R(10)) * (R(1) + fexp(-(u + R(10)) / R(6.24))));
    }
    const R d_inf = R(1) / (R(1) + fexp(-(u + R(10)) / R(8.0)));
    const R tau_f = R(9) / (R(0.0197) * fexp(-R(0.0337 * 0.0337) * ((u + R(10)) * (u + R(10)))) +
                            R(0.02));
    const R f_inf = R(1) / (R(1) + fexp((u + R(28)) / R(6.9)));
    const R tau_fca = R(2);
    const R fca_inf = R(1) / (R(1) + cai / R(0.00035));
    const R d = crn_gate<R>(dr, d_inf, tau_d, dt);
    const R f = crn_gate<R>(fr, f_inf, tau_f, dt);
    const R fca = crn_gate<R>(fcar, fca_inf, tau_fca, dt);
    dr = d;
    fr = f;
    fcar = fca;
    ical = gcal * d * f * fca * (u - R(65));
  }
  // calc_inak
  const R sig = R(1 / 7.0) * (fexp(nao / R(67.3)) - R(1));
  const R fnak = R(1) / (R(1) + R(0.1245) * fexp(R(-0.1) * (F * u) / (R_ * T)) +
                         R(0.0365) * sig * fexp(-(F * u) / (R_ * T)));
  const R inak = inakmax * fnak * (R(1) / (R(1) + fpow(kmnai / nai, R(1.5)))) *
                 (ko / (ko + kmko));
  // calc_inaca
  const R gamma = R(0.35);
  const R exp_term = fexp(gamma * (F * u) / (R_ * T));
  const R exp_rev_term = fexp((gamma - R(1)) * (F * u) / (R_ * T));
  const R nao3 = nao * (nao * nao);
  const R numerator =
      inacamax * (exp_term * (nai * (nai * nai)) * cao - exp_rev_term * nao3 * cai);
  const R term1 = kmnancx * (kmnancx * kmnancx) + nao3;
  const R term2 = kmcancx + cao;
  const R term3 = R(1) + ksatncx * exp_rev_term;
  const R inaca = numerator / (term1 * term2 * term3);
  const R ibca = gcab * (u - eca);
  const R ibna = gnab * (u - ena);
  const R ipca = ipcamax * cai / (cai + R(0.0005));
  // calc_irel (uses the irel stored by the previous step)
  R irel;
  {
    auto&& irr = ref<R>(S, 17, c);
    auto&& vrr = ref<R>(S, 18, c);
    auto&& urr = ref<R>(S, 19, c);
    auto&& wrr = ref<R>(S, 20, c);
    const R tau_u = R(8);
    const R Fn = R(1e-12) * Vrel * irr - R(5 * 1e-13) / F * (R(0.5) * ical - R(0.2) * inaca);
    const R u_inf = R(1) / (R(1) + fexp(-(Fn - R(3.4175e-13)) / R(1.367e-15)));
    const R tau_v = R(1.91) + R(2.09) / (R(1) + fexp(-(Fn - R(3.4175e-13)) / R(1.367e-15)));
    const R v_inf = R(1) - R(1) / (R(1) + fexp(-(Fn - R(6.835e-14)) / R(1.367e-15)));
    R tau_w;
    if constexpr (sizeof(R) == 4) {
      const R y = -(u - R(7.9)) / R(5.0);
      tau_w = R(6.0 / 5.0) * exprel(y) / (R(1) + R(0.3) * fexp(y));
    } else {
      tau_w = R(6) * (R(1) - fexp(-(u - R(7.9)) / R(5.0))) /
              ((R(1) + R(0.3) * fexp(-(u - R(7.9)) / R(5.0))) * (u - R(7.9)));
    }
    const R w_inf = R(1) - R(1) / (R(1) + fexp(-(u - R(40)) / R(17.0)));
    const R urel = crn_gate<R>(urr, u_inf, tau_u, dt);
    const R vrel = crn_gate<R>(vrr, v_inf, tau_v, dt);
    const R wrel = crn_gate<R>(wrr, w_inf, tau_w, dt);
    urr = urel;
    vrr = vrel;
    wrr = wrel;
    irel = krel * (urel * urel) * vrel * wrel * (carel - cai);
    irr = irel;
  }
  const R itr = (caup - carel) / R(180);
  const R iup = iupmax / (R(1) + kup / cai);
  const R iupleak = caup / caupmax * iupmax;
  caupr = caup + dt * (iup - iupleak - itr * (Vrel / Vup));
  nair = nai + dt * ((R(-3) * inak - R(3) * inaca - ibna - ina) / (F * Vj));
  kir = ki + dt * ((R(2) * inak - ik1 - ito - ikur - ikr - iks - ibk) / (F * Vj));
  {
    const R B1 = (R(2) * inaca - ipca - ical - ibca) / (R(2) * F * Vj) +
                 (Vup * (iupleak - iup) + irel * Vrel) / Vj;
    const R B2 = R(1) + trpnmax * kmtrpn / ((cai + kmtrpn) * (cai + kmtrpn)) +
                 cmdnmax * kmcmdn / ((cai + kmcmdn) * (cai + kmcmdn));
    cair = cai + dt * (B1 / B2);
  }
  carelr = carel + dt * ((itr - irel) /
                         (R(1) + csqnmax * kmcsqn / ((carel + kmcsqn) * (carel + kmcsqn))));
  return un - dt * (ina + ik1 + ito + ikur + ikr + iks + ical + ipca + inak +
                    inaca + ibna + ibca);
}

// u and u_new are stored as UType<MODEL, R>: the arithmetic type, except on
// the FP32 path of the millivolt-scale models (LR91, TP06, Courtemanche),
// whose potential is kept double.  There the step's increment dt * sum(I) at
// rest is far below half an ulp of a float near -86 mV (3.8e-6 mV), so a
// float potential stops following the slow relaxation of the resting state,
// and the resting potential of tissue ahead of a front drifts from the FP64
// one (measured 1.6e-3 mV), shifting conduction velocity by ~1e-4 and
// activation times by several steps over a thousand nodes.  With u double the
// currents and gates stay float and only the final u_new - dt * sum(I) is
// double.  The FP64 path is unchanged (U = R = double).
__host__ __device__ constexpr bool u_is_f64(int model, int precision) {
  return precision == 0 || model >= 5;
}

template <int MODEL, typename R>
using UType = std::conditional_t<(sizeof(R) == 8 || MODEL >= 5), double, R>;

template <int MODEL, typename R, typename SP, typename U>
__device__ __forceinline__ U ionic_update(U u, U un, SP& S, int64_t c, R dt,
                                          const Params& P) {
  if constexpr (MODEL == 0) return ionic_ap<R>(u, un, S, c, dt, P);
  else if constexpr (MODEL == 1) return ionic_barkley<R>(u, un, S, c, dt, P);
  else if constexpr (MODEL == 2) return ionic_ms<R>(u, un, S, c, dt, P);
  else if constexpr (MODEL == 3) return ionic_fk<R>(u, un, S, c, dt, P);
  else if constexpr (MODEL == 4) return ionic_bo<R>(u, un, S, c, dt, P);
  else if constexpr (MODEL == 5) return ionic_lr91<R, SP, U>(u, un, S, c, dt, P);
  else if constexpr (MODEL == 6) return ionic_tp06<R, SP, U>(u, un, S, c, dt, P);
  else return ionic_crn<R, SP, U>(u, un, S, c, dt, P);
}

// number of state arrays per model (state_vars minus u)
__host__ __device__ constexpr int model_n_state(int model) {
  return model == 0 ? 1 : model == 1 ? 1 : model == 2 ? 1 : model == 3 ? 2
       : model == 4 ? 3 : model == 5 ? 7 : model == 6 ? 19 : 21;
}
// u_new of an updated node.  Every model's update ends in un + X or un - X
// with X free of un.  For the models with long ionic updates (LR91, TP06,
// Courtemanche) the update is formed first with un = 0 and the diffusion
// result (`diffusion()`, evaluated afterwards) is added last, so the
// neighbour values and the diffusion result do not stay live across the
// ionic code; the light models keep the literal order, which hides the
// neighbour loads better.  Without contraction (FP64) both orders give the
// reference's value exactly.  FP32 builds fuse multiply-adds within one
// expression only, so the single-tissue and batched kernels, which both use
// this function, stay bit-identical.
template <int MODEL, typename R, typename U, typename SP, typename Diffusion>
__device__ __forceinline__ U node_update(U uc, SP& S, int64_t c, R dt, const Params& P,
                                         Diffusion&& diffusion) {
  if constexpr (MODEL >= 5) {
    const U r = ionic_update<MODEL, R>(uc, U(0), S, c, dt, P);
    return diffusion() + r;
  } else {
    return ionic_update<MODEL, R>(uc, diffusion(), S, c, dt, P);
  }
}

__host__ __device__ constexpr int model_n_params(int model) {
  return model == 0 ? 5 : model == 1 ? 3 : model == 2 ? 5 : model == 3 ? 11
       : model == 4 ? 28 : model == 5 ? 15 : model == 6 ? 49 : 39;
}
// FP64 storage of state k under precision prec (1 = FP32 path)
// TP06 on the FP32 path keeps its calcium concentrations cai, casr, cass
// (k < 3) double; nai, Ki, h and j are float differences (state_offset)
__host__ __device__ constexpr bool state_is_f64(int model, int prec, int k) {
  return prec == 0 ? true : (model == 6 && k < 3);
}

}  // namespace fw
