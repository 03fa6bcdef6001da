// fw_batch_tb.hpp -- the batched step of the light models (state all of the
// arithmetic type, at most three arrays), K steps per launch.
//
// A wavefront owns a unit: 64 consecutive columns of one tissue (one per
// lane) and a run of rows [r0, r1), and marches down the rows.  At row x it
// loads row x's potential, state and node codes (level 0) and advances the
// rows it holds by one level each: level t of row x - t from level t - 1 of
// rows x-t-1 .. x-t+1.  Row x - K leaves at level K and is written.  The
// unit loads K rows above and below its own and K columns either side
// (fewer at the tissue's edges, whose outer layer is never updated) and
// computes them too: level t is exact on nodes at least t nodes inside the
// loaded region, so after K levels its own nodes are exact.  Vertical
// neighbours are in registers, horizontal ones come from the adjacent lanes
// (DPP wavefront shifts): nothing passes through LDS and the wavefronts never
// wait for each other.
//
// Reads and writes go to different arrays, so units never race: the
// potential moves from one buffer of the u/u_new pair to the other and the
// state from `state` to `state_alt` (phase 0) or back (phase 1).  Launches
// come in pairs, phase 0 then 1: 2K steps later the state is back in `state`
// and the potential in the buffer its local step's parity names, the layout
// of the single-step kernel.
//
// Per node and level the arithmetic is the single-step kernel's: the same
// neighbourhood (Nb9), diffuse9 and node_update, and a node that is not
// updated keeps its potential and state, so results are bit for bit those of
// the other launch shapes.  Activation times and the probe are taken from
// levels 0 .. K-1 (the pre-step potentials of the launch's K steps) on the
// unit's own nodes, each node being owned by exactly one unit.
#pragma once

#include <utility>

#include "fw_batch.hpp"

namespace fw {

// The isotropic FP32 flux of one node from its four face coefficients
// c = d * face_w(m, m_opp) (N, W, E, S): flux_iso's operations after its
// coefficient products, which depend on the node codes only and are formed
// once per row.
__device__ __forceinline__ float tb_flux_iso(const float (&c)[4], float uN, float uW, float uE,
                                             float uS, float uc, float s) {
#pragma clang fp contract(off)
  float fl = c[0] * (uN - uc);
  fl = fma(c[1], uW - uc, fl);
  fl = fma(c[2], uE - uc, fl);
  fl = fma(c[3], uS - uc, fl);
  return fma(s, fl, uc);
}

template <int MODEL, typename R, int K, bool AN, bool NT, bool ACT>
__device__ __forceinline__ void tb_march(const BatchArgs<R, R>& a, const BatchTissue& d,
                                         const int32_t ls0, const int32_t phase, const int32_t w0,
                                         const int32_t c0, const int32_t c1, const int32_t r0,
                                         const int32_t r1) {
  constexpr int NS = model_n_state(MODEL);
  constexpr bool EXACT = sizeof(R) == 8;
  // the isotropic FP32 flux from per-row coefficients
  constexpr bool COEF = !AN && !EXACT;
  // Row y's data lives in slot y mod P (relative to the first row marched)
  // from its load at iteration y until its last use: its state until level K
  // (iteration y + K), its node codes until level K of row y + 1 (iteration
  // y + K + 1); a level-l potential is read up to two iterations after it
  // is formed.  P = K + 2 slots keep every live row apart.
  constexpr int P = K + 2;
  static_assert(NS >= 1 && NS <= 3, "light models");
  const int32_t lane = (int32_t)(threadIdx.x & 63);
  const int32_t nj = d.nj, ni = d.ni;
  const int32_t col = w0 + lane;
  const bool in_cols = col < nj;
  const bool own_col = col >= c0 && col < c1;
  // buffers: with K even the pair's first launch writes the buffer of the
  // other parity, its second one reads it back
  const int32_t in_b = (ls0 & 1) ^ ((K % 2 == 0) ? phase : 0);
  const global_t<R>* __restrict__ uin = (const global_t<R>*)(in_b ? d.un : d.u);
  global_t<R>* __restrict__ uout = (global_t<R>*)(in_b ? d.u : d.un);
  const global_t<uint8_t>* __restrict__ code = (const global_t<uint8_t>*)d.code;
  global_t<int32_t>* const act = (global_t<int32_t>*)d.act;
  const global_t<R>* sin[NS];
  global_t<R>* sout[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    sin[k] = (const global_t<R>*)(phase ? d.state_alt[k] : d.state[k]);
    sout[k] = (global_t<R>*)(phase ? d.state[k] : d.state_alt[k]);
  }
  R thr;
  if constexpr (sizeof(R) == 8) thr = d.act_thr;
  else thr = d.act_thr_f;
  double dd[4];
  float df[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    dd[q] = d.dd[q];
    df[q] = d.df[q];
  }
  const int stencil = AN ? FW_STENCIL_ANISO : FW_STENCIL_ISO;
  // activation recorded at local steps >= act_start (ACT: some step of this
  // launch is one); the probe below probe_len, by the unit owning its node
  const int32_t act_start = d.act_start;
  // pi: the probe's row if this unit owns its node, else a value no row
  // index reaches (the march also forms rows above its first, down to -K)
  const int32_t probe_idx = d.probe ? d.probe_idx : -1;
  int32_t pi = probe_idx >= 0 ? probe_idx / nj : INT32_MIN;
  const int32_t pj = probe_idx >= 0 ? probe_idx - pi * nj : -1;
  if (!(pi >= r0 && pi < r1 && pj >= c0 && pj < c1)) pi = INT32_MIN;
  const bool probe_lane = lane == pj - w0;
  global_t<R>* const probe = (global_t<R>*)d.probe;
  const int32_t probe_len = d.probe_len;

  const int32_t x_lo = r0 - K > 0 ? r0 - K : 0;
  const int32_t x_end = r1 + K - 1;  // rows from ni on are outside: code 0

  struct Row {
    R u;
    R s[NS];
    uint32_t code;
    int32_t act;
  };
  // Rows from ni on and columns from nj on lie outside the tissue: they are
  // read at the last row / column (any value serves, their codes are 0 and
  // the values reach no node the unit keeps)
  const int32_t col_c = in_cols ? col : nj - 1;
  auto load = [&](int32_t x) {
    Row r;
    const int32_t c = (x < ni ? x : ni - 1) * nj + col_c;
    r.u = ld_state<NT>(uin + c);
#pragma unroll
    for (int k = 0; k < NS; ++k) r.s[k] = ld_state<NT>(sin[k] + c);
    const uint32_t cd = code[c];
    r.code = (x < ni && in_cols) ? cd : 0u;
    r.act = INT32_MAX;  // not recorded by this unit: never set, never written
    if constexpr (ACT) {
      if (own_col && x >= r0 && x < r1) r.act = act[c];
    }
    return r;
  };

  R U[K][P];                  // U[l][s]: level-l potential of the row in slot s
  R Uw[AN ? K : 1][P];        // .. and of the columns west and east of it
  R Ue[AN ? K : 1][P];
  uint32_t cw[P];             // node codes: own | west << 8 | east << 16
  float cf[COEF ? P : 1][4];  // face coefficients N, W, E, S
  RegState<R, NS> st[P];
  int32_t A[P];               // activation times (INT32_MAX where not recorded here)
#pragma unroll
  for (int j = 0; j < P; ++j) {
#pragma unroll
    for (int l = 0; l < K; ++l) U[l][j] = R(0);
#pragma unroll
    for (int l = 0; l < (AN ? K : 1); ++l) Uw[l][j] = Ue[l][j] = R(0);
    cw[j] = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) cf[COEF ? j : 0][q] = 0.0f;
#pragma unroll
    for (int k = 0; k < NS; ++k) st[j].v[k] = R(0);
    A[j] = INT32_MAX;
  }
  const float cfN = df[0], cfW = df[1];  // iso: dx on the x faces, dy on the y faces

  auto set_u = [&](int l, int j, R v) {
    U[l][j] = v;
    if constexpr (AN) {
      Uw[l][j] = from_west(v);
      Ue[l][j] = from_east(v);
    }
  };

  Row nxt = load(x_lo);
  int32_t x = x_lo;
  (void)act_start;
  auto body = [&](auto ph) -> bool {
    constexpr int PH = decltype(ph)::value;  // slot of row x
    if (x > x_end) return false;
    const Row cur = nxt;
    nxt = load(x < x_end ? x + 1 : x);  // in flight during this row's levels
    {
      const uint32_t w = (uint32_t)from_west((int32_t)cur.code);
      const uint32_t e = (uint32_t)from_east((int32_t)cur.code);
      cw[PH] = cur.code | (w << 8) | (e << 16);
    }
    if constexpr (COEF) {
      // row x - 1 now has both vertical neighbours' codes
      constexpr int sm = (PH + P - 1) % P, sN = (PH + P - 2) % P;
      const uint32_t pN = cw[sN], pC = cw[sm], pS = cw[PH];
      const int mN = pN & 1, mS = pS & 1, mW = (pC >> 8) & 1, mE = (pC >> 16) & 1;
      cf[sm][0] = cfN * face_w<float>(mN, mS);
      cf[sm][1] = cfW * face_w<float>(mW, mE);
      cf[sm][2] = cfW * face_w<float>(mE, mW);
      cf[sm][3] = cfN * face_w<float>(mS, mN);
    }
    set_u(0, PH, cur.u);
#pragma unroll
    for (int k = 0; k < NS; ++k) st[PH].v[k] = cur.s[k];
    A[PH] = cur.act;
    // times set in this launch lie in [ls0, ls0 + K); earlier ones below ls0
    if (ACT && ls0 >= act_start) A[PH] = (A[PH] < 0 && cur.u > thr) ? ls0 : A[PH];
    if (x == pi && ls0 < probe_len) {
      if (probe_lane) probe[ls0] = cur.u;
    }
    // levels 1 .. K: row x - t from level t - 1 of rows x-t-1 .. x-t+1
#pragma unroll
    for (int t = 1; t <= K; ++t) {
      const int sy = (PH + 2 * P - t) % P, sN = (PH + 2 * P - t - 1) % P,
                sS = (PH + 2 * P - t + 1) % P;
      const int l = t - 1;
      const R uc = U[l][sy];
      const uint32_t pC = cw[sy];
      const RegState<R, NS> s0 = st[sy];
      RegState<R, NS> s1 = s0;
      R val;
      if constexpr (COEF) {
        const float uW = from_west(uc), uE = from_east(uc);
        val = node_update<MODEL, R>(uc, s1, 0, a.dt, a.P, [&] {
          return tb_flux_iso(cf[sy], U[l][sN], uW, uE, U[l][sS], uc, a.s);
        });
      } else {
        const uint32_t pN = cw[sN], pS = cw[sS];
        Nb9<R> q;
        q.v[4] = uc;
        q.v[1] = U[l][sN];
        q.v[7] = U[l][sS];
        if constexpr (AN) {
          q.v[3] = Uw[l][sy];
          q.v[5] = Ue[l][sy];
          q.v[0] = Uw[l][sN];
          q.v[2] = Ue[l][sN];
          q.v[6] = Uw[l][sS];
          q.v[8] = Ue[l][sS];
        } else {
          q.v[3] = from_west(uc);
          q.v[5] = from_east(uc);
          q.v[0] = q.v[2] = q.v[6] = q.v[8] = R(0);
        }
        q.m[0] = (pN >> 8) & 1;
        q.m[1] = pN & 1;
        q.m[2] = (pN >> 16) & 1;
        q.m[3] = (pC >> 8) & 1;
        q.m[4] = pC & 1;
        q.m[5] = (pC >> 16) & 1;
        q.m[6] = (pS >> 8) & 1;
        q.m[7] = pS & 1;
        q.m[8] = (pS >> 16) & 1;
        val = node_update<MODEL, R>(uc, s1, 0, a.dt, a.P, [&] {
          return diffuse9<R, EXACT>(stencil, dd, df, q, a.D, a.dtd, a.dr2, a.inv_dr2, a.s);
        });
      }
      const bool upd = (pC & 3) == 1;
      const R nv = upd ? val : uc;
#pragma unroll
      for (int k = 0; k < NS; ++k) st[sy].v[k] = upd ? s1.v[k] : s0.v[k];
      const int32_t y = x - t;
      if (t < K) {
        set_u(t, sy, nv);
        if (ACT && ls0 + t >= act_start) A[sy] = (A[sy] < 0 && nv > thr) ? ls0 + t : A[sy];
        if (y == pi && ls0 + t < probe_len) {
          if (probe_lane) probe[ls0 + t] = nv;
        }
      } else if (own_col && y >= r0 && y < r1) {
        const int32_t c = y * nj + col;
        st_state<NT>(uout + c, nv);
#pragma unroll
        for (int k = 0; k < NS; ++k) st_state<NT>(sout[k] + c, st[sy].v[k]);
      }
    }
    {
      // row x - K + 1 has passed its last check
      constexpr int sa = (PH + 2 * P - (K - 1)) % P;
      const int32_t ya = x - K + 1;
      if (ACT && ya >= r0 && ya < r1 && own_col && A[sa] >= ls0 && A[sa] < ls0 + K)
        act[ya * nj + col] = A[sa];
    }
    ++x;
    return true;
  };
  while (tb_phases(body, std::make_integer_sequence<int, P>{})) {
  }
}

template <int MODEL, typename R, int K, bool AN, bool NT>
__global__ __launch_bounds__(kBlock) void k_batch_tb(const BatchArgs<R, R> a, const TbLaunch L) {
  const uint32_t b = blockIdx.x;
  const uint32_t lb = (b % kXcds) * L.per_xcd + b / kXcds;
  const uint32_t unit = __builtin_amdgcn_readfirstlane(lb * kTbWaves + (threadIdx.x >> 6));
  const int32_t step = a.step_in ? *a.step_in : a.step;
  if (a.step_out && b == 0 && threadIdx.x == 0) *a.step_out = step + K;
  if (lb >= L.nblocks || unit >= L.n_units) return;  // whole wavefronts leave
  const TbUnit u = L.units[unit];
  const int32_t t = __builtin_amdgcn_readfirstlane(u.tissue);
  const int32_t w0 = __builtin_amdgcn_readfirstlane(u.w0);
  const int32_t cols = __builtin_amdgcn_readfirstlane(u.cols);
  const int32_t rows = __builtin_amdgcn_readfirstlane(u.rows);
  const BatchTissue& d = a.T[t];
  const int32_t ls0 = step - d.step0;
  if (ls0 < 0) return;  // not admitted yet: the host launches no such unit
  const int32_t c0 = cols & 0xffff, c1 = (int32_t)((uint32_t)cols >> 16);
  const int32_t r0 = rows & 0xffff, r1 = (int32_t)((uint32_t)rows >> 16);
  if (d.act && d.act_start >= 0 && ls0 + K - 1 >= d.act_start)
    tb_march<MODEL, R, K, AN, NT, true>(a, d, ls0, L.phase, w0, c0, c1, r0, r1);
  else
    tb_march<MODEL, R, K, AN, NT, false>(a, d, ls0, L.phase, w0, c0, c1, r0, r1);
}

}  // namespace fw
