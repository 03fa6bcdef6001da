// fw_batch_inst.hip -- instantiates the batched step for one (precision,
// model) pair, selected by FW_PREC and FW_MODEL like fw_step_inst.hip.
#include <type_traits>

#include "fw_batch.hpp"
#include "fw_internal.hpp"

#ifndef FW_PREC
#error "FW_PREC must be defined"
#endif
#ifndef FW_MODEL
#error "FW_MODEL must be defined"
#endif

#define FW_CAT4_(a, b, c, d) a##b##c##d
#define FW_CAT4(a, b, c, d) FW_CAT4_(a, b, c, d)
#define FW_BLAUNCHER FW_CAT4(fw_batch_launch_p, FW_PREC, _m, FW_MODEL)

namespace fw {

using BReal = std::conditional_t<FW_PREC == 0, double, float>;
using UReal = UType<FW_MODEL, BReal>;  // storage type of u / u_new

int FW_BLAUNCHER(const BatchLaunch& L, hipStream_t s) {
  BatchArgs<BReal, UReal> a;
  a.T = static_cast<const BatchTissue*>(L.table);
  a.btis = L.btis;
  a.bstart = L.bstart;
  a.nblocks = L.nblocks;
  a.per_xcd = L.per_xcd;
  a.step_in = L.step_in;
  a.step_out = L.step_out;
  a.step = L.step;
  a.D = L.D;
  a.dtd = L.dt;
  a.dr2 = L.dr2;
  a.inv_dr2 = L.inv_dr2;
  a.dt = (BReal)L.dt;
  a.s = (UReal)(L.D * L.dt / L.dr2);
  a.P = L.P;
  const dim3 gx(L.per_xcd * kXcds), g1(L.nblocks), blk(kBlock);
  switch (L.variant) {
#define FW_BCASE(V)                                                                          \
  case V:                                                                                    \
    hipLaunchKernelGGL((k_batch_step<FW_MODEL, BReal, V>), ((V)&kBatchXcd) ? gx : g1, blk, 0, \
                       s, a);                                                                \
    break;
    FW_BCASE(0)
    FW_BCASE(kBatchXcd)
    FW_BCASE(kBatchNT)
    FW_BCASE(kBatchXcd | kBatchNT)
#if FW_MODEL <= 4
    FW_BCASE(kBatchVec)
    FW_BCASE(kBatchVec | kBatchXcd)
    FW_BCASE(kBatchVec | kBatchNT)
    FW_BCASE(kBatchVec | kBatchXcd | kBatchNT)
#endif
#undef FW_BCASE
    default:
      set_error("batch launch: variant %d not available for model %d", L.variant, FW_MODEL);
      return FW_E_INVALID;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("batch launch failed: %s", hipGetErrorString(e));
    return FW_E_HIP;
  }
  return FW_OK;
}

}  // namespace fw
