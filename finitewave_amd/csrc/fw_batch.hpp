// fw_batch.hpp -- many 2D tissues stepped by one launch per time step.
//
// Each tissue keeps its own HBM buffers (u/u_new pair, state arrays, node
// codes, activation-time map, probe trace) and a record in a device table.
// A block table maps every block of the launch to a run of consecutive
// nodes of one tissue, so tissues of uneven size share a launch without
// padding, and the host admits or retires tissues by rewriting the block
// table between launches.  Fields are uniform per tissue (scalar
// conductivity, one fibre direction), held as four half-step values in the
// record.  Each tissue counts its own local step (global step - step0),
// which selects its u/u_new orientation, gates its activation-time
// recording and indexes its probe trace.
//
// The kernel is shaped for a memory-bound step:
//  * the tissue record is read with uniform (scalar) loads at entry, before
//    any store; node indices are 32-bit; no node coordinate is divided out;
//  * for the light models (state all of the arithmetic type, at most three
//    arrays) a work-item steps 4 consecutive nodes: each row of the 3x6
//    neighbourhood is one 4-wide vector load plus its two edge values, the
//    node codes of a row one 32-bit load, the state one vector load per
//    array, and the results one vector store per array -- a quarter of the
//    memory instructions of one node per work-item.  Nodes that are not
//    updated are written back unchanged, so every store covers whole lines;
//  * the per-node arithmetic is one function (diffuse9 + ionic_update) for
//    both layouts, the same operations in the same order as the
//    single-tissue kernel, so results do not depend on the layout.
// Admission checks that no updated node (code & 3 == 1) lies on the
// tissue's outer layer and that no node is frozen (code bit 1): the kernel
// then needs no column guard, and rewriting a non-updated node's u_new with
// its u is exact (such a node is never stimulated or updated, so both
// buffers hold its initial value).
#pragma once

#include "../../include/finitewave_amd.h"
#include "fw_kernels.hpp"

namespace fw {

// device record of one tissue, built from the public fw_tissue_desc
struct BatchTissue {
  void* u;
  void* un;
  const uint8_t* code;
  int32_t* act;
  void* probe;
  void* state[FW_MAX_STATE];
  void* state_alt[FW_BATCH_ALT_STATE];  // the multi-step shapes' second state set
  const int32_t* nodes;  // the node-list shapes: updated nodes and the probe node, ascending
  double dd[4];      // half-step fields (FP64 weights)
  double act_thr;
  float df[4];       // the same fields rounded to float (FP32 flux form)
  float act_thr_f;
  int32_t n, ni, nj, step0, act_start, probe_idx, probe_len, stencil, n_nodes;
};

// A unit of the multi-step launch shape (fw_batch_tb.hpp): tissue, first
// column of its 64-lane window, own columns [c0, c1) and own rows [r0, r1)
// (16 bits each).
struct alignas(16) TbUnit {
  int32_t tissue;
  int32_t w0;
  int32_t cols;  // c0 | c1 << 16
  int32_t rows;  // r0 | r1 << 16
};

constexpr int kTbWaves = kBlock / 64;  // units per block

struct TbLaunch {
  const TbUnit* __restrict__ units;  // this launch's units (of one stencil)
  uint32_t n_units;
  uint32_t nblocks;  // launched blocks hold kBlock / 64 units each
  uint32_t per_xcd;  // blocks per XCD of the XCD-contiguous map
  int32_t phase;     // 0: state -> state_alt, 1: state_alt -> state
  int32_t aniso;     // the units' tissues use the 9-point stencil
};

// R: arithmetic type of the ionic update; U: storage type of u / u_new
template <typename R, typename U = R>
struct BatchArgs {
  const BatchTissue* __restrict__ T;
  const int32_t* __restrict__ btis;    // tissue of each logical block
  const int32_t* __restrict__ bstart;  // first local node of each logical block
  uint32_t nblocks, per_xcd;
  const int32_t* step_in;
  int32_t* step_out;
  int32_t step;
  double D, dtd, dr2, inv_dr2;
  R dt;
  U s;  // D * dt / dr**2 in the potential's type
  Params P;
};

// 3x3 neighbourhood of one node, row-major: NW N NE / W C E / SW S SE
// (rows are i-1, i, i+1; columns j-1, j, j+1): potentials and conduct bits
template <typename R>
using Nb9 = Hood2<R>;

// diffused u at an interior node of a uniform-field tissue: FP64 exact
// weights (the reference's weight arithmetic and summation order, EXACT) or
// the flux form evaluated in R, the potential's type (double fields when R is
// double, their float roundings when R is float -- the single-tissue
// kernel's operands)
template <typename R, bool EXACT>
__device__ __forceinline__ R diffuse9(int stencil, const double* dd, const float* df,
                                      const Nb9<R>& q, double D, double dtd, double dr2,
                                      double inv_dr2, R s) {
  const R uc = q.v[4];
  if (stencil == FW_STENCIL_ISO) {
    const int mN = q.m[1], mS = q.m[7], mW = q.m[3], mE = q.m[5];
    if constexpr (EXACT) {
      double W[5];
      W[0] = iso_comp(dd[0], mN, mS);
      W[1] = iso_comp(dd[1], mW, mE);
      W[3] = iso_comp(dd[1], mE, mW);
      W[4] = iso_comp(dd[0], mS, mN);
      W[2] = -(W[0] + W[1] + W[3] + W[4]);
      scale_exact<5>(W, 2, D, dtd, dr2, inv_dr2);
      return q.v[1] * W[0] + q.v[3] * W[1] + uc * W[2] + q.v[5] * W[3] + q.v[7] * W[4];
    } else {
      const R dx = sizeof(R) == 8 ? (R)dd[0] : (R)df[0];
      const R dy = sizeof(R) == 8 ? (R)dd[1] : (R)df[1];
      return flux_iso<R>(q, dx, dy, dy, dx, s);
    }
  }
  Nbr M;
#pragma unroll
  for (int k = 0; k < 27; ++k) M.m[k] = 0;
#pragma unroll
  for (int di = 0; di < 3; ++di)
#pragma unroll
    for (int dj = 0; dj < 3; ++dj) M.m[di * 9 + dj * 3 + 1] = q.m[di * 3 + dj];
  if constexpr (EXACT) {
    const double dxx = dd[0], dxy = dd[1], dyx = dd[2], dyy = dd[3];
    double W[9];
    wts_aniso2d(W, M, dxx, dxx, dxy, dxy, dyx, dyx, dyy, dyy);
    scale_exact<9>(W, 4, D, dtd, dr2, inv_dr2);
    return q.v[0] * W[0] + q.v[1] * W[1] + q.v[2] * W[2] + q.v[3] * W[3] + uc * W[4] +
           q.v[5] * W[5] + q.v[6] * W[6] + q.v[7] * W[7] + q.v[8] * W[8];
  } else {
    const R dxx = sizeof(R) == 8 ? (R)dd[0] : (R)df[0];
    const R dxy = sizeof(R) == 8 ? (R)dd[1] : (R)df[1];
    const R dyx = sizeof(R) == 8 ? (R)dd[2] : (R)df[2];
    const R dyy = sizeof(R) == 8 ? (R)dd[3] : (R)df[3];
    return flux_aniso<R>(q, dxx, dxx, dxy, dxy, dyx, dyx, dyy, dyy, s);
  }
}

// Launch-shape variants of the batched step (chosen per batch, e.g. by the
// autotuner of finitewave_amd.batch; all give the same results bit for bit):
//   bit 0: 4 consecutive nodes per work-item (light models only), else 1;
//   bit 1: XCD-contiguous block map (launched block b -> logical block
//          (b mod 8) * per_xcd + b div 8), else launched block = logical block;
//   bit 2: state and new potential read and written with the non-temporal
//          hint (faster once a step's data is well beyond the Infinity Cache).
// A block covers batch_unroll(variant) * kBlock consecutive nodes of a tissue.
//   bit 8 (models 5-7): the single-step shapes over each tissue's list of
//          updated nodes (and its probe node) instead of all its nodes, so
//          no lane of a fibrotic tissue idles through the ionic update;
//   bit 3 (light models): K = batch_tb_k(variant) steps per launch, one
//          wavefront per 64 columns x batch_tb_rows(variant) rows of a tissue
//          (fw_batch_tb.hpp), in pairs of launches; bits 0-2 then give the
//          single-step shape that runs the steps a whole pair does not cover,
//          and bit 2 also the multi-step launch's cache hint.
constexpr int kBatchVec = 1, kBatchXcd = 2, kBatchNT = 4, kBatchVariants = 8;
constexpr int kBatchTB = 8;
constexpr int kBatchList = 256;
constexpr int kBatchTbKs = 2, kBatchTbRowSizes = 2;  // K in {2, 4}, rows in {64, 128}

__host__ __device__ constexpr int batch_tb_k(int variant) { return 2 << ((variant >> 4) & 3); }
__host__ __device__ constexpr int batch_tb_rows(int variant) { return 64 << ((variant >> 6) & 3); }
__host__ __device__ constexpr int batch_tb_variant(int base, int kidx, int ridx) {
  return base | kBatchTB | (kidx << 4) | (ridx << 6);
}

__host__ __device__ constexpr bool batch_base_ok(int model, int base) {
  return base >= 0 && base < kBatchVariants && (model <= 4 || (base & kBatchVec) == 0);
}

__host__ __device__ constexpr bool batch_variant_ok(int model, int variant) {
  if (variant & kBatchList)
    return model >= 5 && (variant & ~kBatchList) < kBatchVariants &&
           batch_base_ok(model, variant & ~kBatchList);
  return (variant & kBatchTB)
             ? (model <= 4 && variant >= 0 && (variant >> 8) == 0 &&
                ((variant >> 4) & 3) < kBatchTbKs && ((variant >> 6) & 3) < kBatchTbRowSizes &&
                batch_base_ok(model, variant & 7))
             : batch_base_ok(model, variant);
}

// the shape without autotuning: for models 5-7 the XCD map without the cache
// hint, which TP06 batches ran fastest with in both precisions (six 400^2
// tissues, FP64: 136.6 against 185 us per step with the hint)
__host__ __device__ constexpr int batch_default_variant(int model) {
  return model <= 4 ? (kBatchVec | kBatchXcd | kBatchNT) : kBatchXcd;
}

__host__ __device__ constexpr int batch_unroll(int variant) {
  return (variant & kBatchVec) ? 4 : 1;
}


// Waves per SIMD the compiler must leave room for in the single-step batch
// kernel.  The FP64 TP06 step (exact weights formed per node) would take 132
// registers (3 waves); held to 4 waves it takes 122 without spills and steps a
// batch of six fibrotic 400^2 tissues in 185 us against 200 (5 waves would
// spill 45).  Other instances keep the compiler's choice.
template <int MODEL, typename R>
constexpr int batch_min_waves() {
  return (MODEL == 6 && sizeof(R) == 8) ? 4 : 1;
}

template <int MODEL, typename R, int V>
__global__ __launch_bounds__(kBlock, (batch_min_waves<MODEL, R>())) void k_batch_step(
    const BatchArgs<R, UType<MODEL, R>> a) {
  using U = UType<MODEL, R>;  // the potential's storage type
  constexpr int KU = batch_unroll(V);
  constexpr bool XCD = (V & kBatchXcd) != 0;
  constexpr bool NT = (V & kBatchNT) != 0;
  constexpr bool EXACT = sizeof(R) == 8;
  constexpr bool LIST = (V & kBatchList) != 0;
  constexpr int NS = model_n_state(MODEL);
  static_assert(KU == 1 || (KU == 4 && MODEL <= 4), "4 nodes per work-item: light models");
  static_assert(!LIST || KU == 1, "node lists: one node per work-item");
  const uint32_t b = blockIdx.x;
  const uint32_t lb = XCD ? (b % kXcds) * a.per_xcd + b / kXcds : b;
  if (lb >= a.nblocks) return;
  // the record, read uniformly before any store
  const int32_t t = __builtin_amdgcn_readfirstlane(a.btis[lb]);
  const int32_t start = __builtin_amdgcn_readfirstlane(a.bstart[lb]);
  const BatchTissue& d = a.T[t];
  const int32_t step = a.step_in ? *a.step_in : a.step;
  const int32_t ls = step - d.step0;
  const int32_t n = d.n, nj = d.nj;
  const bool act_on = d.act && d.act_start >= 0 && ls >= d.act_start;
  const bool probe_on = d.probe && ls < d.probe_len;
  const int32_t probe_idx = d.probe_idx;
  const int stencil = d.stencil;
  // pointers read from memory are generic; typed global, their accesses are
  // global loads and stores, which do not share the scalar loads' counter
  const global_t<uint8_t>* __restrict__ code = (const global_t<uint8_t>*)d.code;
  global_t<U>* const ua = (global_t<U>*)d.u;
  global_t<U>* const ub = (global_t<U>*)d.un;
  global_t<int32_t>* const act = (global_t<int32_t>*)d.act;
  global_t<U>* const probe = (global_t<U>*)d.probe;
  U thr;
  if constexpr (sizeof(U) == 8) thr = d.act_thr;
  else thr = d.act_thr_f;
  double dd[4];
  float df[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    dd[q] = d.dd[q];
    df[q] = d.df[q];
  }
  std::conditional_t<NT, StreamStatePtrs, StatePtrs> S;
#pragma unroll
  for (int k = 0; k < NS; ++k) S.p[k] = d.state[k];

  if (a.step_out && lb == 0 && threadIdx.x == 0) *a.step_out = step + 1;
  if (ls < 0) return;
  const global_t<U>* __restrict__ u = (ls & 1) ? ub : ua;
  global_t<U>* __restrict__ un = (ls & 1) ? ua : ub;
  const bool aniso = stencil != FW_STENCIL_ISO;

  if constexpr (KU == 1) {
    int32_t c = start + (int32_t)threadIdx.x;
    if constexpr (LIST) {
      // blocks run over the list; a listed node is updated unless it is
      // the probe node, which the code check below lets through unchanged
      if (c >= d.n_nodes) return;
      c = d.nodes[c];
    } else {
      if (c >= n) return;
    }
    const uint8_t cd = code[c];
    const U uc = u[c];
    if (act_on && uc > thr && act[c] < 0) act[c] = ls;
    if (probe_on && c == probe_idx) probe[ls] = uc;
    // rows 0 and ni-1 would reach outside the grid
    if ((cd & 3) != 1 || c < nj || c >= n - nj) return;
    Nb9<U> q;
#pragma unroll
    for (int di = -1; di <= 1; ++di)
#pragma unroll
      for (int dj = -1; dj <= 1; ++dj) {
        const int k = (di + 1) * 3 + dj + 1;
        const bool need = aniso || di == 0 || dj == 0;
        q.v[k] = need ? u[c + di * nj + dj] : U(0);
        q.m[k] = need ? (code[c + di * nj + dj] & 1) : 0;
      }
    const U val = node_update<MODEL, R>(uc, S, (int64_t)c, a.dt, a.P, [&] {
      return diffuse9<U, EXACT>(stencil, dd, df, q, a.D, a.dtd, a.dr2, a.inv_dr2, a.s);
    });
    st_state<NT>(un + c, val);
  } else {
    const int32_t c = start + 4 * (int32_t)threadIdx.x;  // a multiple of 4
    if (c >= n) return;
    // the whole 3x6 neighbourhood inside the grid: one vector per row
    if (c >= nj + 1 && c + 4 + nj < n) {
      R uv[3][6];
      int mv[3][6];
      uint32_t cw_c = 0;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int32_t row = c + (r - 1) * nj;
        const Vec4<R> v4 = *(const global_t<Vec4<R>>*)(u + row);
        const uint32_t w = *(const global_t<u32_unaligned>*)(code + row);
        if (r == 1) cw_c = w;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          uv[r][k + 1] = v4[k];
          mv[r][k + 1] = (w >> (8 * k)) & 1;
        }
        const bool edges = aniso || r == 1;
        uv[r][0] = edges ? u[row - 1] : R(0);
        uv[r][5] = edges ? u[row + 4] : R(0);
        mv[r][0] = edges ? (code[row - 1] & 1) : 0;
        mv[r][5] = edges ? (code[row + 4] & 1) : 0;
      }
      if (act_on) {
        Vec4<int32_t> a4 = *(const global_t<Vec4<int32_t>>*)(act + c);
        bool hit = false;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (a4[k] < 0 && uv[1][k + 1] > thr) {
            a4[k] = ls;
            hit = true;
          }
        if (hit) *(global_t<Vec4<int32_t>>*)(act + c) = a4;
      }
      if (probe_on && (uint32_t)(probe_idx - c) < 4u) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (c + k == probe_idx) probe[ls] = uv[1][k + 1];
      }
      Vec4<R> st[NS > 0 ? NS : 1];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        st[s] = ld_state<NT>((const global_t<Vec4<R>>*)((global_t<R>*)S.p[s] + c));
      Vec4<R> out;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        Nb9<R> q;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int dj = 0; dj < 3; ++dj) {
            q.v[r * 3 + dj] = uv[r][k + dj];
            q.m[r * 3 + dj] = mv[r][k + dj];
          }
        const R uc = uv[1][k + 1];
        const bool upd = ((cw_c >> (8 * k)) & 3) == 1;
        RegState<R, NS> rs;
#pragma unroll
        for (int s = 0; s < NS; ++s) rs.v[s] = st[s][k];
        const R val = node_update<MODEL, R>(uc, rs, 0, a.dt, a.P, [&] {
          return diffuse9<R, EXACT>(stencil, dd, df, q, a.D, a.dtd, a.dr2, a.inv_dr2, a.s);
        });
        out[k] = upd ? val : uc;
#pragma unroll
        for (int s = 0; s < NS; ++s) st[s][k] = upd ? rs.v[s] : st[s][k];
      }
      st_state<NT>((global_t<Vec4<R>>*)(un + c), out);
#pragma unroll
      for (int s = 0; s < NS; ++s) st_state<NT>((global_t<Vec4<R>>*)((global_t<R>*)S.p[s] + c), st[s]);
      return;
    }
    // first and last rows and the tissue's tail: node by node
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int32_t cc = c + k;
      if (cc >= n) break;
      const uint8_t cd = code[cc];
      const R uc = u[cc];
      if (act_on && uc > thr && act[cc] < 0) act[cc] = ls;
      if (probe_on && cc == probe_idx) probe[ls] = uc;
      if ((cd & 3) != 1 || cc < nj || cc >= n - nj) continue;
      Nb9<R> q;
#pragma unroll
      for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
          const int kk = (di + 1) * 3 + dj + 1;
          const bool need = aniso || di == 0 || dj == 0;
          q.v[kk] = need ? u[cc + di * nj + dj] : R(0);
          q.m[kk] = need ? (code[cc + di * nj + dj] & 1) : 0;
        }
      RegState<R, NS> rs;
#pragma unroll
      for (int s = 0; s < NS; ++s) rs.v[s] = ld_state<NT>((const global_t<R>*)S.p[s] + cc);
      const R val = node_update<MODEL, R>(uc, rs, 0, a.dt, a.P, [&] {
        return diffuse9<R, EXACT>(stencil, dd, df, q, a.D, a.dtd, a.dr2, a.inv_dr2, a.s);
      });
      st_state<NT>(un + cc, val);
#pragma unroll
      for (int s = 0; s < NS; ++s) st_state<NT>((global_t<R>*)S.p[s] + cc, rs.v[s]);
    }
  }
}

}  // namespace fw
