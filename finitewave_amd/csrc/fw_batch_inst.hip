// fw_batch_inst.hip -- instantiates the batched step for one (precision,
// model) pair, selected by FW_PREC and FW_MODEL like fw_step_inst.hip.
#include <type_traits>

#include "fw_batch.hpp"
#include "fw_internal.hpp"
#if FW_MODEL <= 4
#include "fw_batch_tb.hpp"
#endif

#ifndef FW_PREC
#error "FW_PREC must be defined"
#endif
#ifndef FW_MODEL
#error "FW_MODEL must be defined"
#endif

#define FW_CAT4_(a, b, c, d) a##b##c##d
#define FW_CAT4(a, b, c, d) FW_CAT4_(a, b, c, d)
#define FW_BLAUNCHER FW_CAT4(fw_batch_launch_p, FW_PREC, _m, FW_MODEL)
#define FW_TLAUNCHER FW_CAT4(fw_batch_tb_launch_p, FW_PREC, _m, FW_MODEL)

namespace fw {

using BReal = std::conditional_t<FW_PREC == 0, double, float>;
using UReal = UType<FW_MODEL, BReal>;  // storage type of u / u_new

static BatchArgs<BReal, UReal> kernel_args(const BatchLaunch& L) {
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
  return a;
}

int FW_BLAUNCHER(const BatchLaunch& L, hipStream_t s) {
  const BatchArgs<BReal, UReal> a = kernel_args(L);
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
#else
    FW_BCASE(kBatchList)
    FW_BCASE(kBatchList | kBatchXcd)
    FW_BCASE(kBatchList | kBatchNT)
    FW_BCASE(kBatchList | kBatchXcd | kBatchNT)
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

int FW_TLAUNCHER(const BatchLaunch& L, const TbLaunch& T, hipStream_t s) {
#if FW_MODEL <= 4
  const BatchArgs<BReal, UReal> a = kernel_args(L);
  const dim3 grid(T.per_xcd * kXcds), blk(kBlock);
  const int K = batch_tb_k(L.variant);
  const bool nt = (L.variant & kBatchNT) != 0;
  const bool an = T.aniso != 0;
#define FW_TBCASE(KK, AN, NT)                                                                  \
  if (K == KK && an == AN && nt == NT)                                                         \
    hipLaunchKernelGGL((k_batch_tb<FW_MODEL, BReal, KK, AN, NT>), grid, blk, 0, s, a, T);      \
  else
  FW_TBCASE(2, false, true) FW_TBCASE(2, false, false) FW_TBCASE(2, true, true)
  FW_TBCASE(2, true, false) FW_TBCASE(4, false, true) FW_TBCASE(4, false, false)
  FW_TBCASE(4, true, true) FW_TBCASE(4, true, false) {
    set_error("batch launch: %d steps per launch not built", K);
    return FW_E_UNSUPPORTED;
  }
#undef FW_TBCASE
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("batch launch failed: %s", hipGetErrorString(e));
    return FW_E_HIP;
  }
  return FW_OK;
#else
  (void)L;
  (void)T;
  (void)s;
  set_error("batch launch: model %d has no multi-step shape", FW_MODEL);
  return FW_E_UNSUPPORTED;
#endif
}

}  // namespace fw
