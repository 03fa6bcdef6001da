// fw_kernels.hpp -- gfx950 kernels of the stepping engine.
//
// One fused kernel per time step: each work-item owns one grid node, rebuilds
// its stencil weights from the uint8 node codes (bit 0 = conducting) and the
// half-step diffusion fields, forms the diffusion sum, applies the ionic
// update and writes u_new -- one pass over HBM per step.  Non-updated nodes
// still run the activation-time test (the reference tracker scans every
// node) and point probes.
//
// Two stencil arithmetics:
//  * FP64: the reference's weights exactly -- the same sequence of += / -=
//    into zeroed weights (asymmetric_stencil_2d.py:291-433 and friends), the
//    numpy epilogue ((w*D)*dt)/dr**2 with +1 on the centre, and the
//    left-to-right weighted sum of the diffusion kernel.
//  * FP32: the same operator in flux form, u + s*sum(faces), s = D*dt/dr**2,
//    which avoids the 1 - O(1) cancellation of the centre weight in float.
//
// Grid: one flat 1-D range of nodes, 256 work-items per block, block ids
// remapped so each XCD walks a contiguous slab of rows (neighbour rows then
// hit that XCD's own L2).
#pragma once

#include <utility>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fw_models.hpp"

namespace fw {

struct Field {
  int32_t mode;  // 0 uniform, 1 per-axis table, 2 full
  int32_t axis;
  double value;
  const double* data;
  float fvalue;        // the same in float, for kernels computing in float:
  const float* fdata;  // the plan's float copy of data
};

// How a plan's half-step fields are stored, as far as the step kernel is
// specialised for it: kFieldsUniform, every field uniform; kFieldsK (3D),
// every field a table along the last axis (the plan tabulates uniform ones
// too: layered fibres); kFieldsGeneral, a mode per field, read by dispatch.
enum FieldLayout : int32_t { kFieldsGeneral = 0, kFieldsUniform = 1, kFieldsK = 2 };

// a half-step field at node (i, j, k) = flat index c, in the type F the
// diffusion is computed in; the mode is uniform across a launch
template <typename F>
__device__ __forceinline__ F fval(const Field& f, int64_t i, int64_t j, int64_t k, int64_t c) {
  if (f.mode == 0) {
    if constexpr (sizeof(F) == 4) return f.fvalue;
    else return f.value;
  }
  const int64_t x = f.mode == 1 ? (f.axis == 0 ? i : (f.axis == 1 ? j : k)) : c;
  if constexpr (sizeof(F) == 4) return ((const global_t<float>*)f.fdata)[x];
  else return ((const global_t<double>*)f.data)[x];
}

// a uniform field's value in the type F the diffusion is computed in
template <typename F>
__device__ __forceinline__ F funi(const Field& f) {
  if constexpr (sizeof(F) == 4) return f.fvalue;
  else return f.value;
}

// a field at node c and at c's lower neighbour along axis DIR (flat stride
// `step`): one load when the field is uniform or tabulated along another axis
template <typename F, int DIR>
__device__ __forceinline__ void fval_pair(const Field& f, int64_t i, int64_t j, int64_t k,
                                          int64_t c, int64_t step, F& at_c, F& below) {
  at_c = fval<F>(f, i, j, k, c);
  if (f.mode == 0 || (f.mode == 1 && f.axis != DIR))
    below = at_c;
  else
    below = fval<F>(f, i - (DIR == 0), j - (DIR == 1), k - (DIR == 2), c - step);
}

// fval_pair for a kernel specialised to a field layout FL: uniform values
// straight from the arguments, tables along the last axis indexed by k, or
// the general per-field dispatch
template <typename F, int DIR, int FL>
__device__ __forceinline__ void fval_pair_l(const Field& f, int64_t i, int64_t j, int64_t k,
                                            int64_t c, int64_t step, F& at_c, F& below) {
  if constexpr (FL == kFieldsUniform) {
    at_c = below = funi<F>(f);
  } else if constexpr (FL == kFieldsK) {
    const global_t<F>* t;
    if constexpr (sizeof(F) == 4) t = (const global_t<F>*)f.fdata;
    else t = (const global_t<F>*)f.data;
    at_c = t[k];
    below = DIR == 2 ? t[k - 1] : at_c;
  } else {
    fval_pair<F, DIR>(f, i, j, k, c, step, at_c, below);
  }
}

// Division of 32-bit unsigned numerators by a divisor d >= 1 fixed for a
// launch: q = (t + ((n - t) >> s1)) >> s2 with t = mulhi(n, m), Granlund and
// Montgomery's round-up multiplier (Hacker's Delight, 2nd ed., 10-9), exact
// for every n < 2^32.  Five integer operations instead of the emulated 64-bit
// division; the host computes (m, s1, s2) once per launch.
struct UDiv {
  uint32_t m, s1, s2;
};

inline UDiv make_udiv(uint32_t d) {
  uint32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1;
  return UDiv{(uint32_t)m, l < 1 ? l : 1u, l > 1 ? l - 1 : 0u};
}

__device__ __forceinline__ uint32_t udiv(uint32_t n, UDiv v) {
  const uint32_t t = __umulhi(n, v.m);
  return (t + ((n - t) >> v.s1)) >> v.s2;
}

// The quotients of a cross term of the exact weights for one uniform field
// value d (the counts mh, ml are 3 or 4 whenever the term is on), index
// [count == 4]: nq = -d/ml, pq = d/mh, dc = d * (1/mh - 1/ml).  The host
// fills them with the same IEEE operations (minor_quotients), so a kernel
// for uniform fields selects instead of dividing, with the same bits.
struct MinorQ {
  double nq[2], pq[2], dc[2][2];
};

// One scheduled stimulus (fw_stim_sched): applied to the pre-step potential
// of the steps [s_lo, s_hi] at conducting nodes inside the box [lo, hi) and
// the mask; kind 0 sets `value` (or vals[s - s_lo]), kind 1 adds it and
// clamps at u_max when `clamp`
constexpr int kMaxStims = 8;
struct StimRec {
  int32_t lo[3], hi[3];
  int32_t kind, clamp;
  int32_t s_lo, s_hi;
  const uint8_t* mask;
  const double* vals;
  double value, u_max;
};

// R: arithmetic type of the ionic update; U: storage type of u / u_new
// (UType<MODEL, R>, double for the millivolt-scale models on the FP32 path)
template <typename R, typename U = R>
struct KArgs {
  const U* __restrict__ u;
  U* __restrict__ u_new;
  StatePtrs st;
  const uint8_t* __restrict__ code;
  int64_t ni, nj, nk, N;
  uint32_t nblocks;  // logical blocks covering N
  uint32_t per_xcd;  // logical blocks per XCD
  Field f[9];
  const double* __restrict__ w;  // precomputed weights (stencil 2)
  double D, dtd, dr2, inv_dr2;  // inv_dr2: exact 1/dr2 when dr2 = 2^k, else 0
  R dt;
  U s;  // D * dt / dr**2 in the potential's type
  Params P;
  int32_t* act;
  U act_thr;
  int32_t act_on;
  int32_t step;
  const int32_t* step_in;  // graph replay: step read from / written to device
  int32_t* step_out;
  int32_t n_probe;
  const int64_t* probe_idx;
  const int32_t* probe_var;
  double* probe_buf;
  int32_t probe_row;
  int32_t idx32;      // N < 2^32: node coordinates by 32-bit division
  UDiv div_jk, div_k;  // division by nj * nk and by nk
  int32_t field_layout;  // FieldLayout of the plan's fields
  MinorQ mq[9];          // kFieldsUniform, FP64: each field's cross-term quotients
  // kFieldsUniform, FP64, 2D: the final weights (epilogue applied) of a node
  // whose 3x3 neighbourhood all conducts -- the same for every such node
  double wfull[9];
  // the stimulus schedule: the kernel of step s applies the stimuli of step
  // s + 1 to the u_new it writes while s + 1 < stim_end (*stim_end_in in
  // graph replay)
  int32_t n_stim;
  int32_t stim_end;
  const int32_t* stim_end_in;
  int32_t stim_lo[3], stim_hi[3];  // the union of the entries' boxes
  StimRec stim[kMaxStims];
};

// Lane shifts across the whole 64-lane wavefront by DPP (v_mov_b32_dpp
// wave_shr:1 / wave_shl:1): a VALU move, no LDS round trip.  The bound-control
// bit writes 0 to the lane without a source, so no destination needs
// clearing first.
// lane i receives lane i-1's value (lane 0 receives 0)
template <typename T>
__device__ __forceinline__ T from_west(T x) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(
        T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int32_t, x), 0x138, 0xf, 0xf, true));
  } else {
    const int64_t b = __builtin_bit_cast(int64_t, x);
    const int32_t lo = __builtin_amdgcn_update_dpp(0, (int32_t)b, 0x138, 0xf, 0xf, true);
    const int32_t hi = __builtin_amdgcn_update_dpp(0, (int32_t)(b >> 32), 0x138, 0xf, 0xf, true);
    return __builtin_bit_cast(T, ((int64_t)hi << 32) | (int64_t)(uint32_t)lo);
  }
}

// lane i receives lane i+1's value (lane 63 receives 0)
template <typename T>
__device__ __forceinline__ T from_east(T x) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(
        T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int32_t, x), 0x130, 0xf, 0xf, true));
  } else {
    const int64_t b = __builtin_bit_cast(int64_t, x);
    const int32_t lo = __builtin_amdgcn_update_dpp(0, (int32_t)b, 0x130, 0xf, 0xf, true);
    const int32_t hi = __builtin_amdgcn_update_dpp(0, (int32_t)(b >> 32), 0x130, 0xf, 0xf, true);
    return __builtin_bit_cast(T, ((int64_t)hi << 32) | (int64_t)(uint32_t)lo);
  }
}

// Runs f(integral_constant<0>) .. f(integral_constant<P-1>) while each
// returns true: the body of a loop unrolled P times whose phase is known at
// compile time.
template <typename F, int... I>
__device__ __forceinline__ bool tb_phases(F&& f, std::integer_sequence<int, I...>) {
  return (f(std::integral_constant<int, I>{}) && ...);
}

// (i, j, k) of flat node index c in a C-order grid (k = 0 in 2D)
template <typename A>
__device__ __forceinline__ void node_coords(const A& a, int64_t c, int64_t& i, int64_t& j,
                                            int64_t& k) {
  if (a.idx32) {
    const uint32_t cc = (uint32_t)c;
    const uint32_t ii = udiv(cc, a.div_jk);
    const uint32_t rem = cc - ii * (uint32_t)(a.nj * a.nk);
    const uint32_t jj = udiv(rem, a.div_k);
    i = ii;
    j = jj;
    k = rem - jj * (uint32_t)a.nk;
  } else {
    const int64_t sjk = a.nj * a.nk;
    i = c / sjk;
    const int64_t rem = c - i * sjk;
    j = rem / a.nk;
    k = rem - j * a.nk;
  }
}

// logical block id with each XCD owning a contiguous run of blocks
__device__ __forceinline__ int64_t logical_block(uint32_t per_xcd) {
  const uint32_t b = blockIdx.x;
  return (int64_t)(b % kXcds) * per_xcd + (b / kXcds);
}

// ---------------------------------------------------------------------------
// FP64 exact weights
// ---------------------------------------------------------------------------

// isotropic_stencil_2d.py:135-158: d * m0 * (m0 + (m1 == 0))
__host__ __device__ __forceinline__ double iso_comp(double d, int m0, int m1) {
  return d * (double)m0 * (double)(m0 + (m1 == 0));
}

// asymmetric_stencil_2d.py:204-259 for 0/1 mesh values: zeros unless the two
// face nodes conduct and both 4-node sides hold at least 3 conducting nodes.
// Zeros are returned as +0.0; adding or subtracting a signed zero leaves
// the accumulated weights unchanged, so this matches the reference bits.
// T: d's precomputed quotients, or null to divide.
__host__ __device__ __forceinline__ void minor_exact(double d, const MinorQ* T, int m0, int m1,
                                                     int m2, int m3, int m4, int m5,
                                                     double q[6]) {
  const int mh = m2 + m3 + m4 + m5;
  const int ml = m0 + m1 + m2 + m3;
  if (m2 == 0 || m3 == 0 || mh < 3 || ml < 3) {
    q[0] = q[1] = q[2] = q[3] = q[4] = q[5] = 0.0;
    return;
  }
  if (T) {
    const int h = mh == 4, l = ml == 4;
    const double nq = l ? T->nq[1] : T->nq[0];
    const double pq = h ? T->pq[1] : T->pq[0];
    const double dc = h ? (l ? T->dc[1][1] : T->dc[1][0]) : (l ? T->dc[0][1] : T->dc[0][0]);
    q[0] = m0 ? nq : 0.0;
    q[1] = m1 ? nq : 0.0;
    q[2] = dc;
    q[3] = dc;
    q[4] = m4 ? pq : 0.0;
    q[5] = m5 ? pq : 0.0;
    return;
  }
  const double dmh = (double)mh, dml = (double)ml;
  q[0] = m0 ? (-d) / dml : 0.0;
  q[1] = m1 ? (-d) / dml : 0.0;
  const double c = 1.0 / dmh - 1.0 / dml;  // m2 = m3 = 1 here
  q[2] = d * c;
  q[3] = d * c;
  q[4] = m4 ? d / dmh : 0.0;
  q[5] = m5 ? d / dmh : 0.0;
}

// numpy epilogue: w = w * D * dt / dr**2 ; w[..., centre] += 1.  When dr**2
// is a power of two, inv_dr2 != 0 holds its exact reciprocal and the division
// is the same multiplication (barring subnormal results).
template <int K>
__host__ __device__ __forceinline__ void scale_exact(double W[K], int centre, double D,
                                            double dt, double dr2, double inv_dr2) {
  if (inv_dr2 != 0.0) {
#pragma unroll
    for (int q = 0; q < K; ++q) W[q] = W[q] * D * dt * inv_dr2;
  } else {
#pragma unroll
    for (int q = 0; q < K; ++q) W[q] = W[q] * D * dt / dr2;
  }
  W[centre] += 1.0;
}

// 3x3(x3) neighbourhood of conduct bits, index (di+1)*9 + (dj+1)*3 + (dk+1)
struct Nbr {
  int m[27];
  __host__ __device__ __forceinline__ int operator()(int di, int dj, int dk) const {
    return m[(di + 1) * 9 + (dj + 1) * 3 + (dk + 1)];
  }
};

// asymmetric_stencil_2d.py:291-433 (2D: dk = 0 everywhere)
__host__ __device__ __forceinline__ void wts_aniso2d(double W[9], const Nbr& M,
                                            double dxx0, double dxx1,
                                            double dxy0, double dxy1,
                                            double dyx0, double dyx1,
                                            double dyy0, double dyy1,
                                            const MinorQ* Txy = nullptr,
                                            const MinorQ* Tyx = nullptr) {
#pragma unroll
  for (int q = 0; q < 9; ++q) W[q] = 0.0;
  double q[6], qm;
  qm = dxx0 * (double)M(-1, 0, 0);
  W[1] += qm; W[4] -= qm;
  minor_exact(dxy0, Txy, M(-1, -1, 0), M(0, -1, 0), M(-1, 0, 0), M(0, 0, 0),
              M(-1, 1, 0), M(0, 1, 0), q);
  W[0] -= q[0]; W[3] -= q[1]; W[1] -= q[2]; W[4] -= q[3]; W[2] -= q[4]; W[5] -= q[5];
  qm = dyy0 * (double)M(0, -1, 0);
  W[3] += qm; W[4] -= qm;
  minor_exact(dyx0, Tyx, M(-1, -1, 0), M(-1, 0, 0), M(0, -1, 0), M(0, 0, 0),
              M(1, -1, 0), M(1, 0, 0), q);
  W[0] -= q[0]; W[1] -= q[1]; W[3] -= q[2]; W[4] -= q[3]; W[6] -= q[4]; W[7] -= q[5];
  qm = dyy1 * (double)M(0, 1, 0);
  W[5] += qm; W[4] -= qm;
  minor_exact(dyx1, Tyx, M(-1, 1, 0), M(-1, 0, 0), M(0, 1, 0), M(0, 0, 0),
              M(1, 1, 0), M(1, 0, 0), q);
  W[2] += q[0]; W[1] += q[1]; W[5] += q[2]; W[4] += q[3]; W[8] += q[4]; W[7] += q[5];
  qm = dxx1 * (double)M(1, 0, 0);
  W[7] += qm; W[4] -= qm;
  minor_exact(dxy1, Txy, M(1, -1, 0), M(0, -1, 0), M(1, 0, 0), M(0, 0, 0),
              M(1, 1, 0), M(0, 1, 0), q);
  W[6] += q[0]; W[3] += q[1]; W[7] += q[2]; W[4] += q[3]; W[8] += q[4]; W[5] += q[5];
}

// asymmetric_stencil_3d.py:163-458; d[face][comp]: faces x0 x1 y0 y1 z0 z1,
// comps: x faces (xx, xy, xz), y faces (yy, yx, yz), z faces (zz, zx, zy)
// field of d[face][comp] (the field order xx xy xz yx yy yz zx zy zz)
constexpr int kFaceField[6][3] = {{0, 1, 2}, {0, 1, 2}, {4, 3, 5}, {4, 3, 5}, {8, 6, 7}, {8, 6, 7}};

__device__ __forceinline__ void wts_aniso3d(double W[19], const Nbr& M,
                                            const double d[6][3],
                                            const MinorQ* T9 = nullptr) {
#pragma unroll
  for (int q = 0; q < 19; ++q) W[q] = 0.0;
  double q[6], qm;
  // q (i-1/2, j, k)
  qm = d[0][0] * (double)M(-1, 0, 0);
  W[1] += qm; W[4] -= qm;
  minor_exact(d[0][1], T9 ? T9 + kFaceField[0][1] : nullptr,
              M(-1, -1, 0), M(0, -1, 0), M(-1, 0, 0), M(0, 0, 0),
              M(-1, 1, 0), M(0, 1, 0), q);
  W[0] -= q[0]; W[3] -= q[1]; W[1] -= q[2]; W[4] -= q[3]; W[2] -= q[4]; W[5] -= q[5];
  minor_exact(d[0][2], T9 ? T9 + kFaceField[0][2] : nullptr,
              M(-1, 0, -1), M(0, 0, -1), M(-1, 0, 0), M(0, 0, 0),
              M(-1, 0, 1), M(0, 0, 1), q);
  W[15] -= q[0]; W[11] -= q[1]; W[1] -= q[2]; W[4] -= q[3]; W[17] -= q[4]; W[12] -= q[5];
  // q (i+1/2, j, k)
  qm = d[1][0] * (double)M(1, 0, 0);
  W[7] += qm; W[4] -= qm;
  minor_exact(d[1][1], T9 ? T9 + kFaceField[1][1] : nullptr,
              M(1, -1, 0), M(0, -1, 0), M(1, 0, 0), M(0, 0, 0),
              M(1, 1, 0), M(0, 1, 0), q);
  W[6] += q[0]; W[3] += q[1]; W[7] += q[2]; W[4] += q[3]; W[8] += q[4]; W[5] += q[5];
  minor_exact(d[1][2], T9 ? T9 + kFaceField[1][2] : nullptr,
              M(1, 0, -1), M(0, 0, -1), M(1, 0, 0), M(0, 0, 0),
              M(1, 0, 1), M(0, 0, 1), q);
  W[16] += q[0]; W[11] += q[1]; W[7] += q[2]; W[4] += q[3]; W[18] += q[4]; W[12] += q[5];
  // q (i, j-1/2, k)
  qm = d[2][0] * (double)M(0, -1, 0);
  W[3] += qm; W[4] -= qm;
  minor_exact(d[2][1], T9 ? T9 + kFaceField[2][1] : nullptr,
              M(-1, -1, 0), M(-1, 0, 0), M(0, -1, 0), M(0, 0, 0),
              M(1, -1, 0), M(1, 0, 0), q);
  W[0] -= q[0]; W[1] -= q[1]; W[3] -= q[2]; W[4] -= q[3]; W[6] -= q[4]; W[7] -= q[5];
  minor_exact(d[2][2], T9 ? T9 + kFaceField[2][2] : nullptr,
              M(0, -1, -1), M(0, 0, -1), M(0, -1, 0), M(0, 0, 0),
              M(0, -1, 1), M(0, 0, 1), q);
  W[9] -= q[0]; W[11] -= q[1]; W[3] -= q[2]; W[4] -= q[3]; W[10] -= q[4]; W[12] -= q[5];
  // q (i, j+1/2, k)
  qm = d[3][0] * (double)M(0, 1, 0);
  W[5] += qm; W[4] -= qm;
  minor_exact(d[3][1], T9 ? T9 + kFaceField[3][1] : nullptr,
              M(-1, 1, 0), M(-1, 0, 0), M(0, 1, 0), M(0, 0, 0),
              M(1, 1, 0), M(1, 0, 0), q);
  W[2] += q[0]; W[1] += q[1]; W[5] += q[2]; W[4] += q[3]; W[8] += q[4]; W[7] += q[5];
  minor_exact(d[3][2], T9 ? T9 + kFaceField[3][2] : nullptr,
              M(0, 1, -1), M(0, 0, -1), M(0, 1, 0), M(0, 0, 0),
              M(0, 1, 1), M(0, 0, 1), q);
  W[13] += q[0]; W[11] += q[1]; W[5] += q[2]; W[4] += q[3]; W[14] += q[4]; W[12] += q[5];
  // q (i, j, k-1/2)
  qm = d[4][0] * (double)M(0, 0, -1);
  W[11] += qm; W[4] -= qm;
  minor_exact(d[4][1], T9 ? T9 + kFaceField[4][1] : nullptr,
              M(-1, 0, -1), M(-1, 0, 0), M(0, 0, -1), M(0, 0, 0),
              M(1, 0, -1), M(1, 0, 0), q);
  W[15] -= q[0]; W[1] -= q[1]; W[11] -= q[2]; W[4] -= q[3]; W[16] -= q[4]; W[7] -= q[5];
  minor_exact(d[4][2], T9 ? T9 + kFaceField[4][2] : nullptr,
              M(0, -1, -1), M(0, -1, 0), M(0, 0, -1), M(0, 0, 0),
              M(0, 1, -1), M(0, 1, 0), q);
  W[9] -= q[0]; W[3] -= q[1]; W[11] -= q[2]; W[4] -= q[3]; W[13] -= q[4]; W[5] -= q[5];
  // q (i, j, k+1/2)
  qm = d[5][0] * (double)M(0, 0, 1);
  W[12] += qm; W[4] -= qm;
  minor_exact(d[5][1], T9 ? T9 + kFaceField[5][1] : nullptr,
              M(-1, 0, 1), M(-1, 0, 0), M(0, 0, 1), M(0, 0, 0),
              M(1, 0, 1), M(1, 0, 0), q);
  W[17] += q[0]; W[1] += q[1]; W[12] += q[2]; W[4] += q[3]; W[18] += q[4]; W[7] += q[5];
  minor_exact(d[5][2], T9 ? T9 + kFaceField[5][2] : nullptr,
              M(0, -1, 1), M(0, -1, 0), M(0, 0, 1), M(0, 0, 0),
              M(0, 1, 1), M(0, 1, 0), q);
  W[10] += q[0]; W[3] += q[1]; W[12] += q[2]; W[4] += q[3]; W[14] += q[4]; W[5] += q[5];
}

// symmetric_stencil_2d.py:120-150 (weights builder only)
__device__ __forceinline__ void sym_comp(double dxx, double dxy, double dyx,
                                         double dyy, int m0, int m1, int m2,
                                         int m3, int qx, int qy,
                                         double out[4]) {
  const int msum = m0 + m1 + m2 + m3;
  if (msum < 3) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    return;
  }
  const double qdx = (double)qx * dxx + (double)qy * dyx;
  const double qdy = (double)qx * dxy + (double)qy * dyy;
  const double w0 = (double)(-m0) / (double)(m0 + m1) * qdx -
                    (double)m0 / (double)(m0 + m2) * qdy;
  const double w1 = (double)(-m1) / (double)(m0 + m1) * qdx +
                    (double)m1 / (double)(m1 + m3) * qdy;
  const double w2 = (double)m2 / (double)(m2 + m3) * qdx -
                    (double)m2 / (double)(m0 + m2) * qdy;
  const double w3 = (double)m3 / (double)(m2 + m3) * qdx +
                    (double)m3 / (double)(m1 + m3) * qdy;
  out[0] = 0.5 * w0;
  out[1] = 0.5 * w1;
  out[2] = 0.5 * w2;
  out[3] = 0.5 * w3;
}

// ---------------------------------------------------------------------------
// FP32 flux form
// ---------------------------------------------------------------------------

// (mean of conducting higher-side nodes - mean of lower-side nodes), or 0
// when the face's minor weights are zero (the reference's rule): times d, the
// algebraic value of sum_k q_k u_k for the minor weights q of the reference.
// Callers add it as fl = fma(+-d, minor_diff(...), fl); with the face off
// that adds d * 0 = 0 exactly.  The one multiply-add inside is written as
// fma, with contraction off, so its rounding is the same in every kernel
// that inlines it.
template <typename R>
__device__ __forceinline__ R minor_diff(int m0, R u0, int m1, R u1, int m2, R u2, int m3,
                                        R u3, int m4, R u4, int m5, R u5) {
#pragma clang fp contract(off)
  const int mh = m2 + m3 + m4 + m5;
  const int ml = m0 + m1 + m2 + m3;
  const bool on = m2 != 0 && m3 != 0 && mh >= 3 && ml >= 3;
  const R common = u2 + u3;
  const R sh = common + (m4 ? u4 : R(0)) + (m5 ? u5 : R(0));
  const R sl = common + (m0 ? u0 : R(0)) + (m1 ? u1 : R(0));
  const R ih = mh == 4 ? R(0.25) : R(1.0 / 3.0);
  const R il = ml == 4 ? R(0.25) : R(1.0 / 3.0);
  const R v = fma(sh, ih, -(sl * il));
  return on ? v : R(0);  // a select: no branch around the neighbour loads
}

// d * minor_diff(...) with the select last, the form of the 3D flux: there
// it keeps the 19-point FP32 step at 71 registers (7 waves per SIMD) instead
// of 84 (5).  The 3D step has no batched counterpart to agree with, so it
// lets the compiler fuse across statements (the masked minor flux folds into
// one fma with the running sum).
template <typename R>
__device__ __forceinline__ R minor_flux(R d, int m0, R u0, int m1, R u1, int m2, R u2, int m3,
                                        R u3, int m4, R u4, int m5, R u5) {
#pragma clang fp contract(fast)
  const int mh = m2 + m3 + m4 + m5;
  const int ml = m0 + m1 + m2 + m3;
  const bool on = m2 != 0 && m3 != 0 && mh >= 3 && ml >= 3;
  const R common = u2 + u3;
  const R sh = common + (m4 ? u4 : R(0)) + (m5 ? u5 : R(0));
  const R sl = common + (m0 ? u0 : R(0)) + (m1 ? u1 : R(0));
  const R ih = mh == 4 ? R(0.25) : R(1.0 / 3.0);
  const R il = ml == 4 ? R(0.25) : R(1.0 / 3.0);
  const R v = d * (sh * ih - sl * il);
  return on ? v : R(0);
}

// FP32 face weight factor, branch-free: 0 when the neighbour does not
// conduct (the reference's weight is 0 there), 2 when the opposite one does
// not (its mirror), else 1.  Every neighbour value is loaded unconditionally
// (interior nodes only, so in bounds) and a face weighted 0 adds exactly 0:
// no branch divides the wave around the loads.
template <typename R>
__device__ __forceinline__ R face_w(int m, int m_opp) {
  return (R)(m * (2 - m_opp));
}

// ---------------------------------------------------------------------------
// The fused step
// ---------------------------------------------------------------------------

template <typename R, int MODEL>
__device__ __forceinline__ double state_as_double(const StatePtrs& S, int k,
                                                  int64_t c, int prec) {
  if (state_is_f64(MODEL, prec, k)) return static_cast<const double*>(S.p[k])[c];
  return (double)static_cast<const float*>(S.p[k])[c] + state_offset(MODEL, prec, k);
}

// The 3x3 neighbourhood of an interior 2D node: potentials and conduct bits
// (row-major, centre at 4; the isotropic stencil leaves the corners 0).  The
// flux form takes it as an argument so that the fused step can load it
// before the ionic update and the loads' latency overlaps that arithmetic.
template <typename R>
struct Hood2 {
  R v[9];
  int m[9];
};

template <int ST, typename RA, typename R>
__device__ __forceinline__ Hood2<R> load_hood2(const KArgs<RA, R>& a, int64_t c, R uc) {
  Hood2<R> q;
  const int64_t sN = a.nj;
#pragma unroll
  for (int di = -1; di <= 1; ++di)
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj) {
      const int k = (di + 1) * 3 + dj + 1;
      const bool need = (ST == 1 || di == 0 || dj == 0) && k != 4;
      q.v[k] = need ? a.u[c + di * sN + dj] : R(0);
      q.m[k] = need ? (a.code[c + di * sN + dj] & 1) : 0;
    }
  q.v[4] = uc;
  q.m[4] = 1;
  return q;
}

// The flux-form arithmetic of the 2D stencils, shared by the single-tissue
// and the batched step.  Contraction is off here and the multiply-adds are
// written as fma: the rounding then does not depend on the fusion choices
// the compiler makes in each kernel's context, and both kernels give the
// same bits.  Face values: 0 = the face towards lower i (x) or j (y).
template <typename R>
__device__ __forceinline__ R flux_iso(const Hood2<R>& q, R dN, R dW, R dE, R dS, R s) {
#pragma clang fp contract(off)
  const R uc = q.v[4];
  const int mN = q.m[1], mW = q.m[3], mE = q.m[5], mS = q.m[7];
  R fl = (dN * face_w<R>(mN, mS)) * (q.v[1] - uc);
  fl = fma(dW * face_w<R>(mW, mE), q.v[3] - uc, fl);
  fl = fma(dE * face_w<R>(mE, mW), q.v[5] - uc, fl);
  fl = fma(dS * face_w<R>(mS, mN), q.v[7] - uc, fl);
  return fma(s, fl, uc);
}

template <typename R>
__device__ __forceinline__ R flux_aniso(const Hood2<R>& q, R dxx0, R dxx1, R dxy0, R dxy1,
                                        R dyx0, R dyx1, R dyy0, R dyy1, R s) {
#pragma clang fp contract(off)
  const R uNW = q.v[0], uN = q.v[1], uNE = q.v[2], uW = q.v[3], uc = q.v[4], uE = q.v[5];
  const R uSW = q.v[6], uS = q.v[7], uSE = q.v[8];
  const int mNW = q.m[0], mN = q.m[1], mNE = q.m[2], mW = q.m[3], mE = q.m[5];
  const int mSW = q.m[6], mS = q.m[7], mSE = q.m[8];
  // face x0 (i-1/2): major, then the minor with sign -
  R fl = (dxx0 * (R)mN) * (uN - uc);
  fl = fma(-dxy0, minor_diff<R>(mNW, uNW, mW, uW, mN, uN, 1, uc, mNE, uNE, mE, uE), fl);
  // face y0 (j-1/2)
  fl = fma(dyy0 * (R)mW, uW - uc, fl);
  fl = fma(-dyx0, minor_diff<R>(mNW, uNW, mN, uN, mW, uW, 1, uc, mSW, uSW, mS, uS), fl);
  // face y1 (j+1/2)
  fl = fma(dyy1 * (R)mE, uE - uc, fl);
  fl = fma(dyx1, minor_diff<R>(mNE, uNE, mN, uN, mE, uE, 1, uc, mSE, uSE, mS, uS), fl);
  // face x1 (i+1/2)
  fl = fma(dxx1 * (R)mS, uS - uc, fl);
  fl = fma(dxy1, minor_diff<R>(mSW, uSW, mW, uW, mS, uS, 1, uc, mSE, uSE, mE, uE), fl);
  return fma(s, fl, uc);
}

// FP32 flux form of the 2D isotropic (ST 0) and anisotropic (ST 1) stencils
// on a loaded neighbourhood; R is the potential's storage type.  UNI: every
// field is uniform and its value comes straight from the arguments
template <int ST, bool UNI, typename RA, typename R>
__device__ __forceinline__ R flux2d(const KArgs<RA, R>& a, int64_t i, int64_t j, int64_t c,
                                    R uc, const Hood2<R>& q) {
  const int64_t sN = a.nj;
  if constexpr (ST == 0) {
    R dN, dS, dW, dE;
    if constexpr (UNI) {
      dN = dS = funi<R>(a.f[0]);
      dW = dE = funi<R>(a.f[1]);
    } else {
      fval_pair<R, 0>(a.f[0], i, j, 0, c, sN, dS, dN);
      fval_pair<R, 1>(a.f[1], i, j, 0, c, 1, dE, dW);
    }
    return flux_iso<R>(q, dN, dW, dE, dS, a.s);
  } else {
    R dxx0, dxx1, dxy0, dxy1, dyx0, dyx1, dyy0, dyy1;
    if constexpr (UNI) {
      dxx0 = dxx1 = funi<R>(a.f[0]);
      dxy0 = dxy1 = funi<R>(a.f[1]);
      dyx0 = dyx1 = funi<R>(a.f[2]);
      dyy0 = dyy1 = funi<R>(a.f[3]);
    } else {
      fval_pair<R, 0>(a.f[0], i, j, 0, c, sN, dxx1, dxx0);
      fval_pair<R, 0>(a.f[1], i, j, 0, c, sN, dxy1, dxy0);
      fval_pair<R, 1>(a.f[2], i, j, 0, c, 1, dyx1, dyx0);
      fval_pair<R, 1>(a.f[3], i, j, 0, c, 1, dyy1, dyy0);
    }
    return flux_aniso<R>(q, dxx0, dxx1, dxy0, dxy1, dyx0, dyx1, dyy0, dyy1, a.s);
  }
}

// The FP32 flux form of the 3D anisotropic 19-point stencil at one node, from
// its loaded neighbourhood: conduct bits M, neighbour potentials un18 (index
// (di+1)*9 + (dj+1)*3 + dk+1; corners and centre unused), half-step field
// values d[face][comp] and the centre potential uc.  The 3D step has no
// batched counterpart, so the compiler may fuse across statements here (the
// masked minor flux folds into one fma with the running sum).
template <typename R>
__device__ __forceinline__ R flux3d_aniso(const Nbr& M, const R (&un18)[27], const R (&d)[6][3],
                                          R uc, R s) {
#pragma clang fp contract(fast)
  auto U = [&](int di, int dj, int dk) -> R {
    return M(di, dj, dk) ? un18[(di + 1) * 9 + (dj + 1) * 3 + dk + 1] : R(0);
  };
  R fl = R(0);
  // x0
  fl += (R)d[0][0] * (R)M(-1, 0, 0) * (U(-1, 0, 0) - uc);
  fl -= minor_flux<R>((R)d[0][1], M(-1, -1, 0), U(-1, -1, 0), M(0, -1, 0), U(0, -1, 0),
                      M(-1, 0, 0), U(-1, 0, 0), 1, uc,
                      M(-1, 1, 0), U(-1, 1, 0), M(0, 1, 0), U(0, 1, 0));
  fl -= minor_flux<R>((R)d[0][2], M(-1, 0, -1), U(-1, 0, -1), M(0, 0, -1), U(0, 0, -1),
                      M(-1, 0, 0), U(-1, 0, 0), 1, uc,
                      M(-1, 0, 1), U(-1, 0, 1), M(0, 0, 1), U(0, 0, 1));
  // x1
  fl += (R)d[1][0] * (R)M(1, 0, 0) * (U(1, 0, 0) - uc);
  fl += minor_flux<R>((R)d[1][1], M(1, -1, 0), U(1, -1, 0), M(0, -1, 0), U(0, -1, 0),
                      M(1, 0, 0), U(1, 0, 0), 1, uc,
                      M(1, 1, 0), U(1, 1, 0), M(0, 1, 0), U(0, 1, 0));
  fl += minor_flux<R>((R)d[1][2], M(1, 0, -1), U(1, 0, -1), M(0, 0, -1), U(0, 0, -1),
                      M(1, 0, 0), U(1, 0, 0), 1, uc,
                      M(1, 0, 1), U(1, 0, 1), M(0, 0, 1), U(0, 0, 1));
  // y0
  fl += (R)d[2][0] * (R)M(0, -1, 0) * (U(0, -1, 0) - uc);
  fl -= minor_flux<R>((R)d[2][1], M(-1, -1, 0), U(-1, -1, 0), M(-1, 0, 0), U(-1, 0, 0),
                      M(0, -1, 0), U(0, -1, 0), 1, uc,
                      M(1, -1, 0), U(1, -1, 0), M(1, 0, 0), U(1, 0, 0));
  fl -= minor_flux<R>((R)d[2][2], M(0, -1, -1), U(0, -1, -1), M(0, 0, -1), U(0, 0, -1),
                      M(0, -1, 0), U(0, -1, 0), 1, uc,
                      M(0, -1, 1), U(0, -1, 1), M(0, 0, 1), U(0, 0, 1));
  // y1
  fl += (R)d[3][0] * (R)M(0, 1, 0) * (U(0, 1, 0) - uc);
  fl += minor_flux<R>((R)d[3][1], M(-1, 1, 0), U(-1, 1, 0), M(-1, 0, 0), U(-1, 0, 0),
                      M(0, 1, 0), U(0, 1, 0), 1, uc,
                      M(1, 1, 0), U(1, 1, 0), M(1, 0, 0), U(1, 0, 0));
  fl += minor_flux<R>((R)d[3][2], M(0, 1, -1), U(0, 1, -1), M(0, 0, -1), U(0, 0, -1),
                      M(0, 1, 0), U(0, 1, 0), 1, uc,
                      M(0, 1, 1), U(0, 1, 1), M(0, 0, 1), U(0, 0, 1));
  // z0
  fl += (R)d[4][0] * (R)M(0, 0, -1) * (U(0, 0, -1) - uc);
  fl -= minor_flux<R>((R)d[4][1], M(-1, 0, -1), U(-1, 0, -1), M(-1, 0, 0), U(-1, 0, 0),
                      M(0, 0, -1), U(0, 0, -1), 1, uc,
                      M(1, 0, -1), U(1, 0, -1), M(1, 0, 0), U(1, 0, 0));
  fl -= minor_flux<R>((R)d[4][2], M(0, -1, -1), U(0, -1, -1), M(0, -1, 0), U(0, -1, 0),
                      M(0, 0, -1), U(0, 0, -1), 1, uc,
                      M(0, 1, -1), U(0, 1, -1), M(0, 1, 0), U(0, 1, 0));
  // z1
  fl += (R)d[5][0] * (R)M(0, 0, 1) * (U(0, 0, 1) - uc);
  fl += minor_flux<R>((R)d[5][1], M(-1, 0, 1), U(-1, 0, 1), M(-1, 0, 0), U(-1, 0, 0),
                      M(0, 0, 1), U(0, 0, 1), 1, uc,
                      M(1, 0, 1), U(1, 0, 1), M(1, 0, 0), U(1, 0, 0));
  fl += minor_flux<R>((R)d[5][2], M(0, -1, 1), U(0, -1, 1), M(0, -1, 0), U(0, -1, 0),
                      M(0, 0, 1), U(0, 0, 1), 1, uc,
                      M(0, 1, 1), U(0, 1, 1), M(0, 1, 0), U(0, 1, 0));
  return uc + s * fl;
}

// R here is the potential's storage type (FP32 path of the millivolt-scale
// models: double); EXACT selects the reference's FP64 weights, else the flux
// form evaluated in R; FL: the field layout the caller is specialised for
template <int NDIM, int ST, typename RA, typename R, bool EXACT, int FL = kFieldsGeneral>
__device__ __forceinline__ R diffuse(const KArgs<RA, R>& a, int64_t i, int64_t j,
                                     int64_t k, int64_t c, R uc) {
  const int64_t si = a.nj * a.nk, sj = a.nk;  // 2D: nk = 1, sj = 1, si = nj
  const R* __restrict__ u = a.u;
  if constexpr (NDIM == 2) {
    const int64_t sN = si;  // i stride
    if constexpr (ST == 2) {
      // precomputed (*shape, K) weights summed in the reference kernel's
      // order; K (5 or 9) travels in a.f[0].mode
      const int K = a.f[0].mode;
      const double* W = a.w + c * K;
      if (K == 5)
        return (R)((double)u[c - sN] * W[0] + (double)u[c - 1] * W[1] +
                   (double)uc * W[2] + (double)u[c + 1] * W[3] +
                   (double)u[c + sN] * W[4]);
      return (R)((double)u[c - sN - 1] * W[0] + (double)u[c - sN] * W[1] +
                 (double)u[c - sN + 1] * W[2] + (double)u[c - 1] * W[3] +
                 (double)uc * W[4] + (double)u[c + 1] * W[5] +
                 (double)u[c + sN - 1] * W[6] + (double)u[c + sN] * W[7] +
                 (double)u[c + sN + 1] * W[8]);
    } else if constexpr (!EXACT) {
      const Hood2<R> q = load_hood2<ST>(a, c, uc);
      if constexpr (FL == kFieldsUniform) return flux2d<ST, true>(a, i, j, c, uc, q);
      return a.field_layout == kFieldsUniform ? flux2d<ST, true>(a, i, j, c, uc, q)
                                              : flux2d<ST, false>(a, i, j, c, uc, q);
    } else if constexpr (ST == 0) {
      const int mN = a.code[c - sN] & 1, mS = a.code[c + sN] & 1;
      const int mW = a.code[c - 1] & 1, mE = a.code[c + 1] & 1;
      R dN, dS, dW, dE;
      if constexpr (FL == kFieldsUniform) {
        dN = dS = funi<R>(a.f[0]);
        dW = dE = funi<R>(a.f[1]);
      } else {
        fval_pair<R, 0>(a.f[0], i, j, 0, c, sN, dS, dN);
        fval_pair<R, 1>(a.f[1], i, j, 0, c, 1, dE, dW);
      }
      double W[5];
      if (FL == kFieldsUniform && (mN & mS & mW & mE)) {
        // every neighbour conducts: the host's weights for such a node
#pragma unroll
        for (int q = 0; q < 5; ++q) W[q] = a.wfull[q];
      } else {
        W[0] = iso_comp(dN, mN, mS);
        W[1] = iso_comp(dW, mW, mE);
        W[3] = iso_comp(dE, mE, mW);
        W[4] = iso_comp(dS, mS, mN);
        W[2] = -(W[0] + W[1] + W[3] + W[4]);
        scale_exact<5>(W, 2, a.D, a.dtd, a.dr2, a.inv_dr2);
      }
      return u[c - sN] * W[0] + u[c - 1] * W[1] + uc * W[2] + u[c + 1] * W[3] +
             u[c + sN] * W[4];
    } else {  // anisotropic 9-point, exact weights
      Nbr M;
#pragma unroll
      for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
#pragma unroll
          for (int dk = -1; dk <= 1; ++dk) M.m[(di + 1) * 9 + (dj + 1) * 3 + dk + 1] = 0;
          M.m[(di + 1) * 9 + (dj + 1) * 3 + 1] = a.code[c + di * sN + dj] & 1;
        }
      R dxx0, dxx1, dxy0, dxy1, dyx0, dyx1, dyy0, dyy1;
      if constexpr (FL == kFieldsUniform) {
        dxx0 = dxx1 = funi<R>(a.f[0]);
        dxy0 = dxy1 = funi<R>(a.f[1]);
        dyx0 = dyx1 = funi<R>(a.f[2]);
        dyy0 = dyy1 = funi<R>(a.f[3]);
      } else {
        fval_pair<R, 0>(a.f[0], i, j, 0, c, sN, dxx1, dxx0);
        fval_pair<R, 0>(a.f[1], i, j, 0, c, sN, dxy1, dxy0);
        fval_pair<R, 1>(a.f[2], i, j, 0, c, 1, dyx1, dyx0);
        fval_pair<R, 1>(a.f[3], i, j, 0, c, 1, dyy1, dyy0);
      }
      double W[9];
      int all = 1;
#pragma unroll
      for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) all &= M(di, dj, 0);
      if (FL == kFieldsUniform && all) {
        // the whole 3x3 neighbourhood conducts: the host's weights
#pragma unroll
        for (int q = 0; q < 9; ++q) W[q] = a.wfull[q];
      } else {
        if constexpr (FL == kFieldsUniform)
          wts_aniso2d(W, M, dxx0, dxx1, dxy0, dxy1, dyx0, dyx1, dyy0, dyy1, &a.mq[1], &a.mq[2]);
        else
          wts_aniso2d(W, M, dxx0, dxx1, dxy0, dxy1, dyx0, dyx1, dyy0, dyy1);
        scale_exact<9>(W, 4, a.D, a.dtd, a.dr2, a.inv_dr2);
      }
      return u[c - sN - 1] * W[0] + u[c - sN] * W[1] + u[c - sN + 1] * W[2] +
             u[c - 1] * W[3] + uc * W[4] + u[c + 1] * W[5] +
             u[c + sN - 1] * W[6] + u[c + sN] * W[7] + u[c + sN + 1] * W[8];
    }
  } else {  // NDIM == 3
    if constexpr (ST == 2) {
      const int K = a.f[0].mode;
      const double* W = a.w + c * K;
      if (K == 7)
        return (R)((double)u[c - si] * W[0] + (double)u[c - sj] * W[1] +
                   (double)u[c - 1] * W[2] + (double)uc * W[3] +
                   (double)u[c + 1] * W[4] + (double)u[c + sj] * W[5] +
                   (double)u[c + si] * W[6]);
      return (R)((double)u[c - si - sj] * W[0] + (double)u[c - si] * W[1] +
                 (double)u[c - si + sj] * W[2] + (double)u[c - sj] * W[3] +
                 (double)uc * W[4] + (double)u[c + sj] * W[5] +
                 (double)u[c + si - sj] * W[6] + (double)u[c + si] * W[7] +
                 (double)u[c + si + sj] * W[8] + (double)u[c - sj - 1] * W[9] +
                 (double)u[c - sj + 1] * W[10] + (double)u[c - 1] * W[11] +
                 (double)u[c + 1] * W[12] + (double)u[c + sj - 1] * W[13] +
                 (double)u[c + sj + 1] * W[14] + (double)u[c - si - 1] * W[15] +
                 (double)u[c + si - 1] * W[16] + (double)u[c - si + 1] * W[17] +
                 (double)u[c + si + 1] * W[18]);
    } else if constexpr (ST == 0) {
      const int mXm = a.code[c - si] & 1, mXp = a.code[c + si] & 1;
      const int mYm = a.code[c - sj] & 1, mYp = a.code[c + sj] & 1;
      const int mZm = a.code[c - 1] & 1, mZp = a.code[c + 1] & 1;
      R dXm, dXp, dYm, dYp, dZm, dZp;
      fval_pair_l<R, 0, FL>(a.f[0], i, j, k, c, si, dXp, dXm);
      fval_pair_l<R, 1, FL>(a.f[1], i, j, k, c, sj, dYp, dYm);
      fval_pair_l<R, 2, FL>(a.f[2], i, j, k, c, 1, dZp, dZm);
      if constexpr (EXACT) {
        double W[7];
        W[0] = iso_comp(dXm, mXm, mXp);
        W[1] = iso_comp(dYm, mYm, mYp);
        W[2] = iso_comp(dZm, mZm, mZp);
        W[4] = iso_comp(dZp, mZp, mZm);
        W[5] = iso_comp(dYp, mYp, mYm);
        W[6] = iso_comp(dXp, mXp, mXm);
        W[3] = -(W[0] + W[1] + W[2] + W[4] + W[5] + W[6]);
        scale_exact<7>(W, 3, a.D, a.dtd, a.dr2, a.inv_dr2);
        return u[c - si] * W[0] + u[c - sj] * W[1] + u[c - 1] * W[2] + uc * W[3] +
               u[c + 1] * W[4] + u[c + sj] * W[5] + u[c + si] * W[6];
      } else {
        const R uXm = u[c - si], uYm = u[c - sj], uZm = u[c - 1];
        const R uZp = u[c + 1], uYp = u[c + sj], uXp = u[c + si];
        R fl = R(0);
        fl += (R)dXm * face_w<R>(mXm, mXp) * (uXm - uc);
        fl += (R)dYm * face_w<R>(mYm, mYp) * (uYm - uc);
        fl += (R)dZm * face_w<R>(mZm, mZp) * (uZm - uc);
        fl += (R)dZp * face_w<R>(mZp, mZm) * (uZp - uc);
        fl += (R)dYp * face_w<R>(mYp, mYm) * (uYp - uc);
        fl += (R)dXp * face_w<R>(mXp, mXm) * (uXp - uc);
        return uc + a.s * fl;
      }
    } else {  // anisotropic 19-point
      Nbr M;
#pragma unroll
      for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
          for (int dk = -1; dk <= 1; ++dk) {
            const int corner = (di != 0) && (dj != 0) && (dk != 0);
            M.m[(di + 1) * 9 + (dj + 1) * 3 + dk + 1] =
                corner ? 0 : (a.code[c + di * si + dj * sj + dk] & 1);
          }
      // fields: 0 xx 1 xy 2 xz 3 yx 4 yy 5 yz 6 zx 7 zy 8 zz
      R d[6][3];
      fval_pair_l<R, 0, FL>(a.f[0], i, j, k, c, si, d[1][0], d[0][0]);
      fval_pair_l<R, 0, FL>(a.f[1], i, j, k, c, si, d[1][1], d[0][1]);
      fval_pair_l<R, 0, FL>(a.f[2], i, j, k, c, si, d[1][2], d[0][2]);
      fval_pair_l<R, 1, FL>(a.f[4], i, j, k, c, sj, d[3][0], d[2][0]);
      fval_pair_l<R, 1, FL>(a.f[3], i, j, k, c, sj, d[3][1], d[2][1]);
      fval_pair_l<R, 1, FL>(a.f[5], i, j, k, c, sj, d[3][2], d[2][2]);
      fval_pair_l<R, 2, FL>(a.f[8], i, j, k, c, 1, d[5][0], d[4][0]);
      fval_pair_l<R, 2, FL>(a.f[6], i, j, k, c, 1, d[5][1], d[4][1]);
      fval_pair_l<R, 2, FL>(a.f[7], i, j, k, c, 1, d[5][2], d[4][2]);
      if constexpr (EXACT) {
        double W[19];
        wts_aniso3d(W, M, d, FL == kFieldsUniform ? a.mq : nullptr);
        scale_exact<19>(W, 4, a.D, a.dtd, a.dr2, a.inv_dr2);
        return u[c - si - sj] * W[0] + u[c - si] * W[1] + u[c - si + sj] * W[2] +
               u[c - sj] * W[3] + uc * W[4] + u[c + sj] * W[5] +
               u[c + si - sj] * W[6] + u[c + si] * W[7] + u[c + si + sj] * W[8] +
               u[c - sj - 1] * W[9] + u[c - sj + 1] * W[10] + u[c - 1] * W[11] +
               u[c + 1] * W[12] + u[c + sj - 1] * W[13] + u[c + sj + 1] * W[14] +
               u[c - si - 1] * W[15] + u[c + si - 1] * W[16] +
               u[c - si + 1] * W[17] + u[c + si + 1] * W[18];
      } else {
#pragma clang fp contract(fast)
        // the 18 neighbours, loaded unconditionally; U() masks them
        R un18[27];
#pragma unroll
        for (int di = -1; di <= 1; ++di)
#pragma unroll
          for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
            for (int dk = -1; dk <= 1; ++dk) {
              const int q = (di + 1) * 9 + (dj + 1) * 3 + dk + 1;
              const int corner = (di != 0) && (dj != 0) && (dk != 0);
              un18[q] = (corner || q == 13) ? R(0) : u[c + di * si + dj * sj + dk];
            }
        return flux3d_aniso<R>(M, un18, d, uc, a.s);
      }
    }
  }
}

// Waves per SIMD the compiler must leave room for.  The 2D TP06 steps hold 5:
// FP64 then takes 96 registers and spills 7 (uniform-field instance; 1026^2,
// 132.3 -> 127.4 us per step against 4 waves without spills; the compiler
// alone would choose 3); FP32 takes 91-94 without spills.  The 3D
// anisotropic FP64 step would spill 74 registers under a 128-register
// budget.  Other steps keep the compiler's choice.
template <int NDIM, int MODEL, int PREC>
constexpr int step_min_waves() {
  return (NDIM == 2 && MODEL == 6) ? 5 : 1;
}

// 4 values of consecutive nodes, element-aligned: one vector load or store
template <typename T>
struct Vec4T;
template <>
struct Vec4T<float> {
  typedef float type __attribute__((ext_vector_type(4))) __attribute__((aligned(4)));
};
template <>
struct Vec4T<double> {
  typedef double type __attribute__((ext_vector_type(4))) __attribute__((aligned(8)));
};
template <>
struct Vec4T<int32_t> {
  typedef int32_t type __attribute__((ext_vector_type(4))) __attribute__((aligned(4)));
};
template <typename T>
using Vec4 = typename Vec4T<T>::type;
typedef uint32_t __attribute__((aligned(1))) u32_unaligned;

// A point probe's value at node c (pre-step potential uc, post-update state)
template <typename R, int MODEL, int PREC, typename A>
__device__ __forceinline__ void probe_node(const A& a, int64_t c, uint8_t code, double uc) {
  if ((code & 4) && a.probe_row >= 0) {
    for (int p = 0; p < a.n_probe; ++p) {
      if (a.probe_idx[p] != c) continue;
      const int var = a.probe_var[p];
      const double val = var < 0 ? uc : state_as_double<R, MODEL>(a.st, var, c, PREC);
      a.probe_buf[(int64_t)a.probe_row * a.n_probe + p] = val;
    }
  }
}

// Whether any scheduled stimulus applies at step s (uniform across a launch)
template <typename A>
__device__ __forceinline__ bool stim_due(const A& a, int32_t s) {
  if (a.n_stim == 0) return false;
  const int32_t end = a.stim_end_in ? *a.stim_end_in : a.stim_end;
  if (s >= end) return false;
  bool due = false;
  for (int q = 0; q < a.n_stim; ++q) due |= s >= a.stim[q].s_lo && s <= a.stim[q].s_hi;
  return due;
}

// The scheduled stimuli of step s applied, in schedule order, to the value v
// a conducting node (i, j, k) = c holds before step s: the operations of
// Stim*.stimulate on that node (u = V; u += dt*I, then u_max), in the
// potential's storage type as fw_stim_apply does them
// Whether node (i, j, k) lies in the union of the scheduled stimuli's boxes:
// most nodes fail this test and skip the per-entry loop
template <typename A>
__device__ __forceinline__ bool in_stim_box(const A& a, int64_t i, int64_t j, int64_t k) {
  return i >= a.stim_lo[0] && i < a.stim_hi[0] && j >= a.stim_lo[1] && j < a.stim_hi[1] &&
         k >= a.stim_lo[2] && k < a.stim_hi[2];
}

template <typename U, typename A>
__device__ __forceinline__ U stim_node(const A& a, int32_t s, int64_t i, int64_t j, int64_t k,
                                       int64_t c, U v) {
  for (int q = 0; q < a.n_stim; ++q) {
    const StimRec& r = a.stim[q];
    if (s < r.s_lo || s > r.s_hi) continue;
    if (i < r.lo[0] || i >= r.hi[0] || j < r.lo[1] || j >= r.hi[1] || k < r.lo[2] ||
        k >= r.hi[2])
      continue;
    if (r.mask && r.mask[c] == 0) continue;
    if (r.kind == 0) {
      v = (U)(r.vals ? r.vals[s - r.s_lo] : r.value);
    } else {
      v = v + (U)r.value;
      if (r.clamp && v > (U)r.u_max) v = (U)r.u_max;
    }
  }
  return v;
}

// One node of the fused step: activation time, the update of an interior
// updated node, probes.  FL: the plan's field layout (KArgs::field_layout).
// The uniform and along-k instances read the field values without the
// per-field mode dispatch, whose dependent argument loads stalled the waves
// (1026^2 TP06 FP32: 42 -> 40 us per step)
template <int NDIM, int ST, int MODEL, typename R, int PREC, int FL>
__device__ __forceinline__ void step_node(const KArgs<R, UType<MODEL, R>>& a, int64_t c,
                                          int32_t step) {
  using U = UType<MODEL, R>;
  const uint8_t code = a.code[c];
  const U uc = a.u[c];
  if (a.act_on) {
    if (uc > a.act_thr && a.act[c] < 0) a.act[c] = step;
  }
  if ((code & 3) == 1) {
    // node coordinates; skipped work for nodes on the outer layer keeps every
    // neighbour access inside the grid whatever the codes say
    int64_t i, j, k;
    node_coords(a, c, i, j, k);
    bool interior = i > 0 && i < a.ni - 1 && j > 0 && j < a.nj - 1;
    if constexpr (NDIM == 3) interior = interior && k > 0 && k < a.nk - 1;
    if (interior) {
      U v;
      if constexpr (NDIM == 2 && PREC != 0 && ST != 2) {
        // FP32 flux form: the neighbourhood is loaded before the ionic
        // update, as in the batched step
        const Hood2<U> q = load_hood2<ST>(a, c, uc);
        v = node_update<MODEL, R>(uc, a.st, c, a.dt, a.P, [&] {
          return flux2d<ST, FL == kFieldsUniform>(a, i, j, c, uc, q);
        });
      } else {
        v = node_update<MODEL, R>(uc, a.st, c, a.dt, a.P, [&] {
          return diffuse<NDIM, ST, R, U, PREC == 0, FL>(a, i, j, k, c, uc);
        });
      }
      // the next step's stimuli, on the potential this step hands it; the
      // coordinates are recomputed so they need not live across the update
      if (stim_due(a, step + 1)) {
        int64_t si, sj, sk;
        node_coords(a, c, si, sj, sk);
        if (in_stim_box(a, si, sj, sk)) v = stim_node<U>(a, step + 1, si, sj, sk, c, v);
      }
      a.u_new[c] = v;
    }
  } else if ((code & 3) == 3 && stim_due(a, step + 1)) {
    // a frozen node is never updated; the next step's stimuli still apply
    // to the value its u_new buffer holds (the next step's potential)
    int64_t i, j, k;
    node_coords(a, c, i, j, k);
    if (in_stim_box(a, i, j, k)) a.u_new[c] = stim_node<U>(a, step + 1, i, j, k, c, a.u_new[c]);
  }
  probe_node<R, MODEL, PREC>(a, c, code, (double)uc);
}

template <int NDIM, int ST, int MODEL, typename R, int PREC, int FL>
__global__ __launch_bounds__(kBlock, (step_min_waves<NDIM, MODEL, PREC>())) void k_step(
    const KArgs<R, UType<MODEL, R>> a) {
  const int64_t lb = logical_block(a.per_xcd);
  if (lb >= a.nblocks) return;
  const int64_t c = lb * kBlock + threadIdx.x;
  if (c >= a.N) return;
  const int32_t step = a.step_in ? *a.step_in : a.step;
  step_node<NDIM, ST, MODEL, R, PREC, FL>(a, c, step);
  if (a.step_out && c == 0) *a.step_out = step + 1;
}

// ECG2DTracker / ECG3DTracker (ecg_2d_tracker.py:61-151,
// ecg_3d_tracker.py:100-140): per measuring point e, the sum over updated
// nodes of (u_tr - u) / (d_e * dr), u_tr the diffused pre-step potential (the
// fused step's diffusion value before the ionic update) and d_e the squared
// index distance (2D: its third coordinate squared added).  Each block sums
// its nodes with wavefront shuffles and LDS, then adds its partial sums to
// out[e] with one FP64 atomic per point.
constexpr int kEcgPoints = 8;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NDIM, int ST, int MODEL, typename R, int PREC>
__global__ __launch_bounds__(kBlock) void k_ecg(const KArgs<R, UType<MODEL, R>> a,
                                                const double* __restrict__ coords,
                                                int32_t n_points, double dr,
                                                double* __restrict__ out) {
  using U = UType<MODEL, R>;
  __shared__ double part[kBlock / 64][kEcgPoints];
  const int64_t lb = logical_block(a.per_xcd);
  if (lb >= a.nblocks) return;
  const int64_t c = lb * kBlock + threadIdx.x;
  double acc[kEcgPoints];
#pragma unroll
  for (int e = 0; e < kEcgPoints; ++e) acc[e] = 0.0;
  if (c < a.N && (a.code[c] & 3) == 1) {
    int64_t i, j, k;
    node_coords(a, c, i, j, k);
    bool interior = i > 0 && i < a.ni - 1 && j > 0 && j < a.nj - 1;
    if constexpr (NDIM == 3) interior = interior && k > 0 && k < a.nk - 1;
    if (interior) {
      const U uc = a.u[c];
      const U val = diffuse<NDIM, ST, R, U, PREC == 0>(a, i, j, k, c, uc);
      const double du = (double)val - (double)uc;
#pragma unroll
      for (int e = 0; e < kEcgPoints; ++e) {
        if (e >= n_points) break;
        const double x = coords[3 * e], y = coords[3 * e + 1], z = coords[3 * e + 2];
        const double zz = NDIM == 3 ? z - (double)k : z;
        const double d = (x - (double)i) * (x - (double)i) + (y - (double)j) * (y - (double)j) +
                         zz * zz;
        if (d > 0) acc[e] = du / (d * dr);
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < kEcgPoints; ++e) {
    if (e >= n_points) break;
    const double w = wave_sum(acc[e]);
    if (lane == 0) part[wave][e] = w;
  }
  __syncthreads();
  if (threadIdx.x < n_points) {
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) sum += part[w][threadIdx.x];
    atomicAdd(out + threadIdx.x, sum);
  }
}

// ---------------------------------------------------------------------------
// Literal sweeps (one numba kernel each)
// ---------------------------------------------------------------------------

// compute_weights over all nodes; d arrays are full fields, m int8 0/1
struct DPtrs {
  const double* p[9];
};

template <int NDIM, int KIND>
__global__ __launch_bounds__(kBlock) void k_weights(double* __restrict__ w,
                                                    const int8_t* __restrict__ m,
                                                    const DPtrs dp, int64_t ni,
                                                    int64_t nj, int64_t nk) {
  const double* const* d = dp.p;
  const int64_t N = ni * nj * nk;
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= N) return;
  if (m[c] != 1) return;
  const int64_t sjk = nj * nk, si = sjk, sj = nk;
  const int64_t i = c / sjk, j = (c - i * sjk) / nk, k = c - i * sjk - j * nk;
  if (i == 0 || i == ni - 1 || j == 0 || j == nj - 1) return;
  if (NDIM == 3 && (k == 0 || k == nk - 1)) return;
  if constexpr (NDIM == 2 && KIND == 0) {
    const int mN = m[c - si], mS = m[c + si], mW = m[c - 1], mE = m[c + 1];
    double* W = w + c * 5;
    W[0] = iso_comp(d[0][c - si], mN, mS);
    W[1] = iso_comp(d[1][c - 1], mW, mE);
    W[3] = iso_comp(d[1][c], mE, mW);
    W[4] = iso_comp(d[0][c], mS, mN);
    W[2] = -(W[0] + W[1] + W[3] + W[4]);
  } else if constexpr (NDIM == 3 && KIND == 0) {
    double* W = w + c * 7;
    W[0] = iso_comp(d[0][c - si], m[c - si], m[c + si]);
    W[1] = iso_comp(d[1][c - sj], m[c - sj], m[c + sj]);
    W[2] = iso_comp(d[2][c - 1], m[c - 1], m[c + 1]);
    W[4] = iso_comp(d[2][c], m[c + 1], m[c - 1]);
    W[5] = iso_comp(d[1][c], m[c + sj], m[c - sj]);
    W[6] = iso_comp(d[0][c], m[c + si], m[c - si]);
    W[3] = -(W[0] + W[1] + W[2] + W[4] + W[5] + W[6]);
  } else if constexpr (NDIM == 2 && KIND == 1) {
    Nbr M;
    for (int q = 0; q < 27; ++q) M.m[q] = 0;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj) M.m[(di + 1) * 9 + (dj + 1) * 3 + 1] = m[c + di * si + dj];
    double W[9];
    wts_aniso2d(W, M, d[0][c - si], d[0][c], d[1][c - si], d[1][c], d[2][c - 1], d[2][c],
                d[3][c - 1], d[3][c]);
    for (int q = 0; q < 9; ++q) w[c * 9 + q] = W[q];
  } else if constexpr (NDIM == 3 && KIND == 1) {
    Nbr M;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = -1; dk <= 1; ++dk) {
          const int corner = (di != 0) && (dj != 0) && (dk != 0);
          M.m[(di + 1) * 9 + (dj + 1) * 3 + dk + 1] =
              corner ? 0 : m[c + di * si + dj * sj + dk];
        }
    double dd[6][3];
    dd[0][0] = d[0][c - si]; dd[0][1] = d[1][c - si]; dd[0][2] = d[2][c - si];
    dd[1][0] = d[0][c];      dd[1][1] = d[1][c];      dd[1][2] = d[2][c];
    dd[2][0] = d[4][c - sj]; dd[2][1] = d[3][c - sj]; dd[2][2] = d[5][c - sj];
    dd[3][0] = d[4][c];      dd[3][1] = d[3][c];      dd[3][2] = d[5][c];
    dd[4][0] = d[8][c - 1];  dd[4][1] = d[6][c - 1];  dd[4][2] = d[7][c - 1];
    dd[5][0] = d[8][c];      dd[5][1] = d[6][c];      dd[5][2] = d[7][c];
    double W[19];
    wts_aniso3d(W, M, dd);
    for (int q = 0; q < 19; ++q) w[c * 19 + q] = W[q];
  } else if constexpr (NDIM == 2 && KIND == 3) {
    // symmetric_stencil_2d.py:159-250, corner fields at (i-1/2, j-1/2) etc.
    double* W = w + c * 9;
    double o[4];
    int64_t p = c - si - 1;
    sym_comp(d[0][p], d[1][p], d[2][p], d[3][p], m[c - si - 1], m[c - si], m[c - 1], m[c], -1,
             -1, o);
    W[0] += o[0]; W[1] += o[1]; W[3] += o[2]; W[4] += o[3];
    p = c - si;
    sym_comp(d[0][p], d[1][p], d[2][p], d[3][p], m[c - si], m[c - si + 1], m[c], m[c + 1], -1,
             1, o);
    W[1] += o[0]; W[2] += o[1]; W[4] += o[2]; W[5] += o[3];
    p = c - 1;
    sym_comp(d[0][p], d[1][p], d[2][p], d[3][p], m[c - 1], m[c], m[c + si - 1], m[c + si], 1,
             -1, o);
    W[3] += o[0]; W[4] += o[1]; W[6] += o[2]; W[7] += o[3];
    p = c;
    sym_comp(d[0][p], d[1][p], d[2][p], d[3][p], m[c], m[c + 1], m[c + si], m[c + si + 1], 1,
             1, o);
    W[4] += o[0]; W[5] += o[1]; W[7] += o[2]; W[8] += o[3];
  }
}

template <int NDIM, int K>
__global__ __launch_bounds__(kBlock) void k_diffusion(double* __restrict__ u_new,
                                                      const double* __restrict__ u,
                                                      const double* __restrict__ w,
                                                      const int64_t* __restrict__ idx,
                                                      int64_t n, int64_t nj, int64_t nk) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= n) return;
  const int64_t c = idx[q];
  const double* W = w + c * K;
  if constexpr (NDIM == 2 && K == 5) {
    u_new[c] = u[c - nj] * W[0] + u[c - 1] * W[1] + u[c] * W[2] + u[c + 1] * W[3] +
               u[c + nj] * W[4];
  } else if constexpr (NDIM == 2 && K == 9) {
    u_new[c] = u[c - nj - 1] * W[0] + u[c - nj] * W[1] + u[c - nj + 1] * W[2] +
               u[c - 1] * W[3] + u[c] * W[4] + u[c + 1] * W[5] + u[c + nj - 1] * W[6] +
               u[c + nj] * W[7] + u[c + nj + 1] * W[8];
  } else if constexpr (NDIM == 3 && K == 7) {
    const int64_t si = nj * nk, sj = nk;
    u_new[c] = u[c - si] * W[0] + u[c - sj] * W[1] + u[c - 1] * W[2] + u[c] * W[3] +
               u[c + 1] * W[4] + u[c + sj] * W[5] + u[c + si] * W[6];
  } else {
    const int64_t si = nj * nk, sj = nk;
    u_new[c] = u[c - si - sj] * W[0] + u[c - si] * W[1] + u[c - si + sj] * W[2] +
               u[c - sj] * W[3] + u[c] * W[4] + u[c + sj] * W[5] + u[c + si - sj] * W[6] +
               u[c + si] * W[7] + u[c + si + sj] * W[8] + u[c - sj - 1] * W[9] +
               u[c - sj + 1] * W[10] + u[c - 1] * W[11] + u[c + 1] * W[12] +
               u[c + sj - 1] * W[13] + u[c + sj + 1] * W[14] + u[c - si - 1] * W[15] +
               u[c + si - 1] * W[16] + u[c - si + 1] * W[17] + u[c + si + 1] * W[18];
  }
}

template <int MODEL, typename R>
__global__ __launch_bounds__(kBlock) void k_ionic(UType<MODEL, R>* __restrict__ u_new,
                                                  const UType<MODEL, R>* __restrict__ u,
                                                  const StatePtrs st,
                                                  const int64_t* __restrict__ idx,
                                                  int64_t n, R dt, const Params P) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= n) return;
  const int64_t c = idx[q];
  u_new[c] = ionic_update<MODEL, R>(u[c], u_new[c], st, c, dt, P);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_stim(R* __restrict__ u,
                                                 const uint8_t* __restrict__ code,
                                                 const uint8_t* __restrict__ mask,
                                                 int64_t lo0, int64_t lo1, int64_t lo2,
                                                 int64_t b0, int64_t b1, int64_t b2,
                                                 int64_t nj, int64_t nk, int kind,
                                                 double value, const double* value_dev,
                                                 int clamp, double u_max) {
  const int64_t n = b0 * b1 * b2;
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= n) return;
  const int64_t a0 = q / (b1 * b2), r = q - a0 * b1 * b2, a1 = r / b2, a2 = r - a1 * b2;
  const int64_t c = ((lo0 + a0) * nj + (lo1 + a1)) * nk + (lo2 + a2);
  if ((code[c] & 1) == 0) return;
  if (mask && mask[c] == 0) return;
  if (kind == 0) {
    u[c] = (R)(value_dev ? *value_dev : value);
  } else {
    R v = u[c] + (R)value;
    if (clamp && v > (R)u_max) v = (R)u_max;
    u[c] = v;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_act(int32_t* __restrict__ act,
                                                const R* __restrict__ u, int64_t n,
                                                R thr, int32_t step) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= n) return;
  if (act[c] < 0 && u[c] > thr) act[c] = step;
}

}  // namespace fw
