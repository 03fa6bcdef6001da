// fw_march3d.hpp -- the FP32 19-point step of 3D plans whose half-step fields
// are uniform or tables along k (layered fibres), for the two- to
// four-variable models: each face flux and each cross-term square computed
// once instead of by every node that touches it.
//
// The flux form of the anisotropic 19-point stencil (asymmetric_stencil_3d.py
// :163-458 in flux form): the flux into a node c across its upper face along
// axis a is
//   F_a(c) = d_aa * m(c) m(c+a) * (u(c+a) - u(c))
//          + sum over the two other axes b of d_ab * [on] * (Q(S_ab(c)) - Q(S_ab(c-b)))
// where S_ab(n) is the 2x2 square {n, n+a, n+b, n+a+b}, Q its mean over
// conducting nodes, and [on] requires c and c+a to conduct and both squares
// to hold at least three conducting nodes (the reference's minor_component,
// asymmetric_stencil_2d.py:246-259).  The flux across the lower face is
// -F_a(c-a), and the node update is u + s * sum of the six face fluxes.
//
// Work decomposition.  A wavefront takes 64 consecutive nodes of one (j, k)
// plane in C order (p = j*nk + k); lanes 1..63 own their node, lane 0 holds
// the node before them, whose upper-z quantities lane 1 needs.  A block of
// four wavefronts marches along i over a chunk of planes, starting one plane
// early (a halo plane whose results are not written).  Per plane a lane
// loads plane i+1 at rows j-1, j, j+1 and columns k, k+1, computes five
// squares and the fluxes across its upper x, y and z faces and across the
// upper y face of the node below it in j; the lower-x quantities come from
// the previous plane (registers), the lower-z ones from lane - 1 (a DPP
// wavefront shift).
// The face fluxes are the reference operator's; their rounding differs from
// the per-node kernel's (k_step) by the order of the sums.
#pragma once

#include "fw_kernels.hpp"

namespace fw {

constexpr int kMarchOwn = 63;    // nodes a wavefront owns (lanes 1..63)
constexpr int kMarchWaves = 4;   // wavefronts per block
constexpr int kMarchChunk = 16;  // planes a block owns along i

struct MarchGeom {
  int64_t plane;     // nj * nk
  int32_t tiles;     // wavefront tiles per plane: ceil(plane / kMarchOwn)
  int32_t groups;    // blocks per plane: ceil(tiles / kMarchWaves)
  int32_t chunks;    // blocks along i: ceil(ni / kMarchChunk)
  uint32_t nblocks;  // groups * chunks
  uint32_t per_xcd;  // logical blocks per XCD
};

// masked sum and count of a 2x2 square of nodes (counts as floats: 0..4)
struct Sq {
  float s, m;
};

__device__ __forceinline__ Sq square4(float m0, float u0, float m1, float u1, float m2, float u2,
                                      float m3, float u3) {
#pragma clang fp contract(fast)
  Sq q;
  q.s = m0 * u0 + m1 * u1 + m2 * u2 + m3 * u3;
  q.m = (m0 + m1) + (m2 + m3);
  return q;
}

// the square's mean over its conducting nodes (3 or 4 of them when used)
__device__ __forceinline__ float qmean(Sq q) {
  return q.s * (q.m > 3.5f ? 0.25f : (1.0f / 3.0f));
}

// a square as the cross terms read it: its mean and whether at least three
// of its nodes conduct, formed once per square (each square serves two faces)
struct Qv {
  float q;
  float ok;  // 1 or 0
};

// ok = [m >= 3] for counts m in {0, .., 4}: m - 2 clamped to [0, 1]
__device__ __forceinline__ Qv qv(Sq s) {
  return Qv{qmean(s), fminf(fmaxf(s.m - 2.0f, 0.0f), 1.0f)};
}

// a cross term of the face (c, c+a) between its two squares; mf = m(c) m(c+a).
// The flags are exact 0s and 1s: the product is d (hi.q - lo.q) when the
// face and both squares qualify and 0 otherwise, without a select.
__device__ __forceinline__ float cross_term(float d, float mf, Qv hi, Qv lo) {
#pragma clang fp contract(fast)
  return (d * mf) * (hi.ok * lo.ok) * (hi.q - lo.q);
}

// lane - 1's square by a DPP wavefront shift (lane 0 receives zeros; its
// results are not kept)
__device__ __forceinline__ Qv shfl_up1(Qv a) { return Qv{from_west(a.q), from_west(a.ok)}; }

template <int MODEL, int FL>
__global__ __launch_bounds__(kBlock) void k_step3_march(const KArgs<float, float> a,
                                                        const MarchGeom g) {
#pragma clang fp contract(fast)
  static_assert(kBlock == 64 * kMarchWaves, "one wavefront per 64 lanes");
  const uint32_t b = blockIdx.x;
  const uint32_t lb = (b % kXcds) * g.per_xcd + b / kXcds;
  if (lb >= g.nblocks) return;
  const int32_t group = (int32_t)(lb % (uint32_t)g.groups);
  const int32_t chunk = (int32_t)(lb / (uint32_t)g.groups);
  const int lane = threadIdx.x & 63;
  const int64_t tile = (int64_t)group * kMarchWaves + (threadIdx.x >> 6);
  const int64_t p = tile * kMarchOwn + lane - 1;
  const bool valid = p >= 0 && p < g.plane;
  const bool own = valid && lane > 0 && tile < g.tiles;
  const int64_t pc = p < 0 ? 0 : (p >= g.plane ? g.plane - 1 : p);  // in the plane
  const int64_t j = pc / a.nk, k = pc - j * a.nk;
  // flux work reads rows j-1..j+1 and column k+1: offsets kept inside the grid
  const bool fluxable = valid && j >= 1 && j <= a.nj - 2 && k <= a.nk - 2;
  const int64_t om = fluxable ? -a.nk : 0, op = fluxable ? a.nk : 0, ok = fluxable ? 1 : 0;
  const bool row_interior = j >= 1 && j <= a.nj - 2 && k >= 1 && k <= a.nk - 2;
  // the faces' half-step values: along i and j the same on both sides; the
  // upper z face's at k (the lower one is lane - 1's upper one)
  float dxx, dxy, dxz, dyx, dyy, dyz, dzx, dzy, dzz;
  if constexpr (FL == kFieldsUniform) {
    dxx = a.f[0].fvalue, dxy = a.f[1].fvalue, dxz = a.f[2].fvalue;
    dyx = a.f[3].fvalue, dyy = a.f[4].fvalue, dyz = a.f[5].fvalue;
    dzx = a.f[6].fvalue, dzy = a.f[7].fvalue, dzz = a.f[8].fvalue;
  } else {
    auto t = [&](int q) { return ((const global_t<float>*)a.f[q].fdata)[k]; };
    dxx = t(0), dxy = t(1), dxz = t(2), dyx = t(3), dyy = t(4), dyz = t(5);
    dzx = t(6), dzy = t(7), dzz = t(8);
  }
  const global_t<float>* __restrict__ u = (const global_t<float>*)a.u;
  const global_t<uint8_t>* __restrict__ code = (const global_t<uint8_t>*)a.code;
  const int64_t si = g.plane;
  const int32_t step = a.step_in ? *a.step_in : a.step;
  const bool stim_next = stim_due(a, step + 1);
  const int i0 = chunk * kMarchChunk;
  const int i1 = i0 + kMarchChunk < a.ni ? i0 + kMarchChunk : (int)a.ni;
  const int ibeg = i0 > 0 ? i0 - 1 : 0;
  // plane i at rows j-1 (0), j (1), j+1 (2), columns k and k+1 (p suffix)
  int64_t c = (int64_t)ibeg * si + pc;
  float u0 = u[c + om], u1 = u[c], u2 = u[c + op];
  float u0p = u[c + om + ok], u1p = u[c + ok], u2p = u[c + op + ok];
  uint8_t cd1 = code[c];
  float m0 = (float)(code[c + om] & 1), m1 = (float)(cd1 & 1), m2 = (float)(code[c + op] & 1);
  float m0p = (float)(code[c + om + ok] & 1), m1p = (float)(code[c + ok] & 1);
  float m2p = (float)(code[c + op + ok] & 1);
  // the lower-x quantities: those of plane i - 1
  float fx_lo = 0.0f;
  Qv xy_lo{0.0f, 0.0f}, xym_lo{0.0f, 0.0f}, xz_lo{0.0f, 0.0f};
  for (int i = ibeg; i < i1; ++i, c += si) {
    // plane i + 1; the last plane (i = ni - 1) takes itself again: no node
    // of it is updated and no plane follows, so its fluxes are not used
    const int64_t cn = i + 1 < a.ni ? c + si : c;
    const float n0 = u[cn + om], n1 = u[cn], n2 = u[cn + op];
    const float n0p = u[cn + om + ok], n1p = u[cn + ok], n2p = u[cn + op + ok];
    const uint8_t cn1 = code[cn];
    const float q0 = (float)(code[cn + om] & 1), q1 = (float)(cn1 & 1), q2 = (float)(code[cn + op] & 1);
    const float q0p = (float)(code[cn + om + ok] & 1), q1p = (float)(code[cn + ok] & 1);
    const float q2p = (float)(code[cn + op + ok] & 1);
    // the squares this node starts: S_xy(c), S_xy(c-y), S_xz(c), S_yz(c), S_yz(c-y)
    const Qv xy = qv(square4(m1, u1, q1, n1, m2, u2, q2, n2));
    const Qv xym = qv(square4(m0, u0, q0, n0, m1, u1, q1, n1));
    const Qv xz = qv(square4(m1, u1, q1, n1, m1p, u1p, q1p, n1p));
    const Qv yz = qv(square4(m1, u1, m2, u2, m1p, u1p, m2p, u2p));
    const Qv yzm = qv(square4(m0, u0, m1, u1, m0p, u0p, m1p, u1p));
    // those of the node at k - 1 (lane - 1)
    const Qv xz_z = shfl_up1(xz), yz_z = shfl_up1(yz), yzm_z = shfl_up1(yzm);
    // fluxes across the upper x, y, z faces of c and the upper y face of c - y
    const float mx = m1 * q1, my = m1 * m2, mz = m1 * m1p, my0 = m0 * m1;
    const float fx = dxx * mx * (n1 - u1) + cross_term(dxy, mx, xy, xym) +
                     cross_term(dxz, mx, xz, xz_z);
    const float fy = dyy * my * (u2 - u1) + cross_term(dyx, my, xy, xy_lo) +
                     cross_term(dyz, my, yz, yz_z);
    const float fy0 = dyy * my0 * (u1 - u0) + cross_term(dyx, my0, xym, xym_lo) +
                      cross_term(dyz, my0, yzm, yzm_z);
    const float fz = dzz * mz * (u1p - u1) + cross_term(dzx, mz, xz, xz_lo) +
                     cross_term(dzy, mz, yz, yzm);
    const float fz0 = from_west(fz);
    if (own && i >= i0) {
      const float uc = u1;
      // the activation word is read only for a potential above the
      // threshold: below it nothing can be recorded, and a read whose
      // result the loop waits on costs every plane its latency
      if (a.act_on) {
        if (uc > a.act_thr && a.act[c] < 0) a.act[c] = step;
      }
      if ((cd1 & 3) == 1 && row_interior && i >= 1 && i <= a.ni - 2) {
        const float fl = (fx - fx_lo) + (fy - fy0) + (fz - fz0);
        float v = node_update<MODEL, float>(uc, a.st, c, a.dt, a.P, [&] { return uc + a.s * fl; });
        if (stim_next) v = stim_node<float>(a, step + 1, i, j, k, c, v);
        a.u_new[c] = v;
      } else if ((cd1 & 3) == 3 && stim_next) {
        a.u_new[c] = stim_node<float>(a, step + 1, i, j, k, c, a.u_new[c]);
      }
      probe_node<float, MODEL, 1>(a, c, cd1, (double)uc);
      if (a.step_out && c == 0) *a.step_out = step + 1;
    }
    // plane i + 1 becomes plane i
    fx_lo = fx;
    xy_lo = xy, xym_lo = xym, xz_lo = xz;
    u0 = n0, u1 = n1, u2 = n2, u0p = n0p, u1p = n1p, u2p = n2p;
    m0 = q0, m1 = q1, m2 = q2, m0p = q0p, m1p = q1p, m2p = q2p;
    cd1 = cn1;
  }
}

// the launch geometry of k_step3_march for an ni x nj x nk grid; false when
// the grid's counts do not fit its 32-bit block indexing
inline bool march_geom(int64_t ni, int64_t nj, int64_t nk, MarchGeom& g) {
  g.plane = nj * nk;
  const int64_t tiles = (g.plane + kMarchOwn - 1) / kMarchOwn;
  const int64_t groups = (tiles + kMarchWaves - 1) / kMarchWaves;
  const int64_t chunks = (ni + kMarchChunk - 1) / kMarchChunk;
  const int64_t nb = groups * chunks;
  if (tiles > INT32_MAX || nb > (int64_t)UINT32_MAX / 2) return false;
  g.tiles = (int32_t)tiles;
  g.groups = (int32_t)groups;
  g.chunks = (int32_t)chunks;
  g.nblocks = (uint32_t)nb;
  g.per_xcd = (g.nblocks + kXcds - 1) / kXcds;
  return true;
}

}  // namespace fw
