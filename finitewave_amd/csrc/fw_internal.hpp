// fw_internal.hpp -- declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/finitewave_amd.h"
#include "fw_models.hpp"

struct fw_plan {
  fw_plan_desc desc;
  int64_t N;
  uint32_t nblocks;  // logical blocks of 256 nodes
  uint32_t per_xcd;  // logical blocks per XCD
  uint32_t grid;     // launched blocks = per_xcd * 8
  int32_t K;         // weights per node for FW_STENCIL_WEIGHTS
  int32_t* step_words;  // device int32[4]: graph step words [0..1], stim_end [2]
  // captured step graphs for two buffer orientations (u, u_new) and (u_new,
  // u): per orientation, graphs of 2^b pairs of steps (b = 0..kGraphBits-1),
  // captured when first needed; all keyed on the same tracker and state
  struct GraphSet {
    const void* u;
    const void* un;
    hipGraphExec_t g[6];
    uint32_t last_use;
  } graphs[2];
  uint32_t graph_clock;
  void* graph_state[FW_MAX_STATE];
  int32_t graph_act_on;
  int32_t* graph_act;
  double graph_act_thr;
  int32_t graph_keyed;  // graph_state/act/act_on/act_thr hold a key
  // the stimulus schedule (fw_plan_set_stims), copied into each step's
  // kernel arguments
  int32_t n_stims;
  fw_stim_sched stims[FW_MAX_STIMS];
  // FP32 plans with a float potential: float copies of the tabulated
  // half-step fields (made once by fw_plan_create)
  float* ftab[FW_MAX_FIELDS];
  // the field layout the step kernel is instantiated for (fw::FieldLayout);
  // kFieldsK plans hold every field as a table along the last axis, uniform
  // ones filled with their value, in double and in float
  int32_t field_layout;
  double* ktab[FW_MAX_FIELDS];
  float* fktab[FW_MAX_FIELDS];
};

namespace fw {

struct MinorQ;
// d's cross-term quotients for the uniform-field FP64 step kernels, computed
// on the host without contraction (fw_abi.hip)
void minor_quotients(double d, MinorQ& T);

struct TrackArgs {
  int32_t* act;
  double act_thr;
  int32_t act_on;
  int32_t step;
  const int32_t* step_in;
  int32_t* step_out;
  int32_t n_probe;
  const int64_t* probe_idx;
  const int32_t* probe_var;
  double* probe_buf;
  int32_t probe_row;
  // scheduled stimuli of step + 1 are applied while step + 1 < stim_end,
  // read from *stim_end_in when that is set (graph replay)
  int32_t stim_end;
  const int32_t* stim_end_in;
};

typedef int (*StepLauncher)(const fw_plan* plan, const void* u, void* u_new,
                            const StatePtrs& st, const TrackArgs& tr,
                            hipStream_t stream);

// one launcher per (precision, model), defined in fw_step_inst.hip builds
StepLauncher step_launcher(int precision, int model);

typedef int (*EcgLauncher)(const fw_plan* plan, const void* u, const double* coords,
                           int32_t n_points, double dr, double* out, hipStream_t stream);
EcgLauncher ecg_launcher(int precision, int model);

// device trackers (fw_track.hip)
int launch_lat(int u_f64, const void* u, int32_t* state, int32_t* steps, int32_t cap,
               double thr, int32_t step, int64_t n, int32_t* overflow, hipStream_t s);
int launch_tips(int u_f64, const void* u_prev, const void* u, const int64_t* shape, double thr,
                double* tips, int32_t cap, int32_t* count, hipStream_t s);
int launch_zero_f64(double* p, int64_t n, hipStream_t s);

// Completes a parameter block filled with the model's parameters: appends
// the parameter-only quantities (kTp06Derived for TP06), computed in
// fw_abi.hip, an FP64 unit built without contraction, in the operation order
// of the expressions they replace; then the float copies of every entry.
void complete_params(int model, Params& P);

// literal sweeps (fw_literal.hip)
int launch_weights(int ndim, int kind, double* w, const int8_t* m,
                   const double* const* d_dev, const int64_t* shape,
                   hipStream_t s);
int launch_diffusion(int ndim, int K, double* u_new, const double* u,
                     const double* w, const int64_t* idx, int64_t n,
                     const int64_t* shape, hipStream_t s);
int launch_ionic(int model, int precision, void* u_new, const void* u,
                 const StatePtrs& st, const int64_t* idx, int64_t n,
                 double dt, const Params& P, hipStream_t s);
// value_dev: a device value that replaces stim->value of a voltage stimulus
int launch_stim(int u_f64, void* u, const uint8_t* code,
                const fw_stim_desc* stim, const int64_t* shape, hipStream_t s,
                const double* value_dev = nullptr);
int launch_act(int u_f64, int32_t* act, const void* u, int64_t n,
               double thr, int32_t step, hipStream_t s);

void set_error(const char* fmt, ...);

}  // namespace fw

namespace fw {

struct BatchLaunch {
  const void* table;
  const int32_t* btis;
  const int32_t* bstart;
  uint32_t nblocks, per_xcd;
  const int32_t* step_in;
  int32_t* step_out;
  int32_t step;
  int32_t variant;  // launch shape (fw_batch.hpp: kBatchVec | kBatchXcd | kBatchNT | kBatchList)
  double D, dt, dr2, inv_dr2;
  Params P;
};

typedef int (*BatchLauncher)(const BatchLaunch& L, hipStream_t stream);
BatchLauncher batch_launcher(int precision, int model);

// the multi-step shape of the light models (null for the other models):
// L.variant selects K and the cache hint, T the units and the phase
struct TbLaunch;
typedef int (*BatchTbLauncher)(const BatchLaunch& L, const TbLaunch& T, hipStream_t stream);
BatchTbLauncher batch_tb_launcher(int precision, int model);

}  // namespace fw

struct fw_batch {
  int32_t model, precision, max_tissues;
  double D, dt, dr2, inv_dr2;
  fw::Params P;
  void* table;            // device fw::BatchTissue[max_tissues]
  int32_t* btis;          // device block table
  int32_t* bstart;
  int32_t nblocks, cap_blocks;
  int32_t* step_words;    // device int32[4]: graph step words [0..1], border flag [2]
  int32_t* flags;         // device int32[max_tissues]: per-tissue flags of fw_batch_nonfinite
  hipGraphExec_t graph;       // two steps
  hipGraphExec_t graph_long;  // 2 * kGraphLongPairs steps, for long runs
  int32_t graph_nblocks, graph_variant;
  int32_t variant;        // launch shape of fw_batch_run
  int32_t* active;        // host: the occupied slots listed by fw_batch_set_active
  int32_t n_active;
  fw_tissue_desc* host;   // host mirror of the table
  // multi-step shape: unit table of the active tissues (built for tb_variant)
  void* units;            // device fw::TbUnit[cap_units]
  int32_t n_units, n_units_iso, cap_units, tb_variant;
  hipGraphExec_t tb_graph;       // one pair of multi-step launches
  hipGraphExec_t tb_graph_long;  // kBatchTbLongPairs pairs
  int32_t tb_graph_units, tb_graph_iso, tb_graph_variant;
  // node-list shapes (models 5-7): per slot, the device list of its updated
  // nodes and probe node, its length and the buffer's capacity (entries)
  int32_t** lists;
  int32_t* list_n;
  int64_t* list_cap;
};
