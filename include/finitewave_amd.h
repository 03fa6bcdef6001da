/*
 * finitewave_amd.h -- C ABI of the MI355X (gfx950) stepping engine.
 *
 * Plain C: pointers, sizes, status codes.  Every array argument is a device
 * pointer (HIP), C-order over the grid shape, unless stated otherwise.  No
 * entry point allocates device memory except fw_plan_create (a small
 * parameter block; for an FP32 plan whose potential is float, float copies
 * of its tabulated half-step fields; for a 3D plan whose fields are uniform
 * or tables along the last axis, each field as such a table in double and
 * float -- all read once at creation),
 * fw_batch_create (the tissue descriptor table) and fw_batch_set_active (the
 * block table, when it grows).  `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).  All launches are
 * asynchronous on that stream.
 *
 * Status: 0 on success, negative on error; fw_last_error() returns a
 * thread-local message for the last failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the finitewave source):
 *   fw_compute_weights  <- Stencil.compute_weights's njit loops
 *                          cpuwave2D/stencil/isotropic_stencil_2d.py:161-204,
 *                          cpuwave2D/stencil/asymmetric_stencil_2d.py:291-433,
 *                          cpuwave2D/stencil/symmetric_stencil_2d.py:159-250,
 *                          cpuwave3D/stencil/isotropic_stencil_3d.py:112-149,
 *                          cpuwave3D/stencil/asymmetric_stencil_3d.py:163-458
 *   fw_diffusion        <- diffusion_kernel_{2d,3d}_{iso,aniso}(u_new, u, w,
 *                          indexes): isotropic_stencil_2d.py:97-132,
 *                          asymmetric_stencil_2d.py:164-201,
 *                          isotropic_stencil_3d.py:77-110,
 *                          asymmetric_stencil_3d.py:108-160
 *   fw_ionic            <- ionic_kernel_{2d,3d}(u_new, u, *state, indexes, dt,
 *                          *params) of each model, e.g.
 *                          cpuwave2D/model/aliev_panfilov_2d.py:173-203,
 *                          cpuwave2D/model/tp06_2d.py:987-1103
 *   fw_plan_create/fw_run
 *                       <- the per-step body of CardiacModel.run()
 *                          (core/model/cardiac_model.py:164-178): diffusion +
 *                          ionic + activation-time tracker fused into one
 *                          kernel per step, weights rebuilt from the mesh.
 *   fw_stim_apply       <- Stim*.stimulate (cpuwave2D/stimulation/*.py,
 *                          cpuwave3D/stimulation/*.py)
 *   fw_plan_set_stims / fw_stim_apply_step
 *                       <- StimSequence.stimulate_next + Stim.update_status
 *                          (core/stimulation/stim_sequence.py:65-77,
 *                          core/stimulation/stim.py:52-56) over a run, as a
 *                          device schedule fused into the step kernel
 *   fw_act_time_update  <- ActivationTime2DTracker._track
 *                          (cpuwave2D/tracker/activation_time_2d_tracker.py:61-63)
 *   fw_lat_update       <- LocalActivationTime2DTracker._track
 *                          (local_activation_time_2d_tracker.py:60-102)
 *   fw_spiral_tips      <- _track_tip_line (spiral_wave_core_2d_tracker.py:211-249)
 *   fw_ecg              <- ECG2DTracker.calc_ecg + compute_ecg
 *                          (ecg_2d_tracker.py:61-151)
 *   fw_fibrosis         <- Diffuse2D/3DPattern._generate and
 *                          Structural2D/3DPattern._generate
 *                          (diffuse_2d_pattern.py:86-87,
 *                          structural_2d_pattern.py:97-116): the same
 *                          distribution, drawn on the device
 *   fw_batch_*          <- no counterpart (the reference steps one tissue per
 *                          process); many CardiacModel2D.run() calls of one
 *                          model sharing each step's launch.
 */
#ifndef FINITEWAVE_AMD_H
#define FINITEWAVE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FW_ABI_VERSION 4

/* status codes */
#define FW_OK 0
#define FW_E_INVALID (-1)     /* bad argument */
#define FW_E_UNSUPPORTED (-2) /* combination not built */
#define FW_E_HIP (-3)         /* HIP runtime error */
#define FW_E_NOMEM (-4)       /* allocation failed */

/* ionic models (ids in the order of the reference's model packages) */
#define FW_MODEL_ALIEV_PANFILOV 0
#define FW_MODEL_BARKLEY 1
#define FW_MODEL_MITCHELL_SCHAEFFER 2
#define FW_MODEL_FENTON_KARMA 3
#define FW_MODEL_BUENO_OROVIO 4
#define FW_MODEL_LUO_RUDY91 5
#define FW_MODEL_TP06 6
#define FW_MODEL_COURTEMANCHE 7
#define FW_N_MODELS 8

/* stencils */
#define FW_STENCIL_ISO 0     /* 5-point (2D) / 7-point (3D), weights rebuilt */
#define FW_STENCIL_ANISO 1   /* 9-point (2D) / 19-point (3D), weights rebuilt */
#define FW_STENCIL_WEIGHTS 2 /* precomputed (*shape, K) float64 weights */
#define FW_STENCIL_SYM 3     /* fw_compute_weights only: symmetric 9-point */

/* precision */
#define FW_PREC_FP64 0 /* all state FP64; bit-for-bit the reference order */
#define FW_PREC_FP32 1 /* FP32 state; TP06 cai/casr/cass and the potential of
                          LR91/TP06/Courtemanche kept FP64 (fw_u_is_f64); TP06
                          nai/Ki/h/j stored as differences (fw_state_offset) */

/* half-step diffusion field storage */
#define FW_FIELD_UNIFORM 0 /* one value */
#define FW_FIELD_AXIS 1    /* table over one grid axis (data[shape[axis]]) */
#define FW_FIELD_FULL 2    /* one value per node (data[N]) */

/* node code bits (uint8 per node) */
#define FW_CODE_CONDUCT 1 /* mesh == 1: takes part in the stencil */
#define FW_CODE_FROZEN 2  /* special boundary: conducts, never updated */
#define FW_CODE_PROBE 4   /* sampled by a point-probe tracker */

#define FW_MAX_STATE 24
#define FW_MAX_PARAMS 64
#define FW_MAX_FIELDS 9

typedef struct fw_field {
  int32_t mode; /* FW_FIELD_* */
  int32_t axis; /* for FW_FIELD_AXIS: 0, 1 or 2 */
  double value; /* for FW_FIELD_UNIFORM */
  const double* data; /* device, for FW_FIELD_AXIS / FW_FIELD_FULL */
} fw_field;

/* Description of one tissue simulation (one plan per simulation).
 * fields: half-step diffusion coefficients, d[axis][component] at the face
 * between node n and node n+1 along `axis` (the arrays the reference builds
 * with np.roll): iso 2D {xx, yy}; iso 3D {xx, yy, zz};
 * aniso 2D {xx, xy, yx, yy}; aniso 3D {xx, xy, xz, yx, yy, yz, zx, zy, zz}. */
typedef struct fw_plan_desc {
  int32_t ndim;      /* 2 or 3 */
  int32_t model;     /* FW_MODEL_* */
  int32_t stencil;   /* FW_STENCIL_ISO / ANISO / WEIGHTS */
  int32_t precision; /* FW_PREC_* */
  int64_t shape[3];  /* ni, nj, nk (nk = 1 for 2D) */
  const uint8_t* code; /* device [N] */
  int32_t n_fields;
  fw_field fields[FW_MAX_FIELDS];
  const double* weights; /* device [N*K], FW_STENCIL_WEIGHTS only */
  double D_model, dt, dr2; /* dr2 = dr**2 as the host computed it */
  int32_t n_params;
  double params[FW_MAX_PARAMS]; /* reference ionic_kernel argument order */
} fw_plan_desc;

typedef struct fw_plan fw_plan;

/* Per-run tracking.  act: int32 [N] step index of first crossing, -1 before
 * (NULL: off).  act_flags: host [n_steps], nonzero = the tracker's gate is
 * open on that step.  probes: probe_idx[n_probe] flat node index (device),
 * probe_var[n_probe] (device) -1 = u before the step, k = state k after it;
 * probe_rows: host [n_steps] row of probe_buf (double [rows*n_probe],
 * device) to write on that step, -1 = none. */
typedef struct fw_track_desc {
  int32_t* act;
  double act_thr;
  const uint8_t* act_flags;
  int32_t n_probe;
  const int64_t* probe_idx;
  const int32_t* probe_var;
  double* probe_buf;
  const int32_t* probe_rows;
  /* the plan's scheduled stimuli (fw_plan_set_stims) of every step s + 1 with
   * s + 1 < stim_end are applied by the kernel of step s as it writes u_new;
   * stim_end <= step0 + 1 (e.g. 0) applies none.  The stimuli of the first
   * step of a run are the caller's: fw_stim_apply_step. */
  int64_t stim_end;
} fw_track_desc;

typedef struct fw_stim_desc {
  int32_t kind;       /* 0 = set voltage, 1 = add dt*current */
  int64_t lo[3], hi[3]; /* box [lo, hi) clipped to the grid */
  const uint8_t* mask; /* device [N] or NULL; nonzero = inside */
  double value;       /* voltage, or dt*current already multiplied */
  int32_t clamp;      /* 1: u = min(u, u_max) after adding */
  double u_max;
} fw_stim_desc;

int fw_abi_version(void);
const char* fw_last_error(void);

/* Number of state arrays and parameters of a model, and whether state k is
 * stored as FP64 under `precision` (1) or FP32 (0). */
int fw_model_info(int32_t model, int32_t* n_state, int32_t* n_params);
int fw_state_is_f64(int32_t model, int32_t precision, int32_t k);
/* The value state k is stored relative to under `precision`: a state array
 * holds value - *offset.  0 except on the FP32 path of TP06, which stores
 * nai, Ki, h and j (k = 3, 4, 6, 7) as float differences from their initial
 * values 7.67, 138.3, 0.75, 0.75 (a float holding the value itself stops
 * following their slow changes).  Applies to every entry point that takes
 * state arrays (fw_ionic, fw_run, batches); probes record values. */
int fw_state_offset(int32_t model, int32_t precision, int32_t k, double* offset);
/* Whether u / u_new are FP64 (1) or FP32 (0) under `precision`: FP64 on the
 * FP64 path, and on the FP32 path for LR91, TP06 and Courtemanche, whose
 * resting potential a float cannot follow (increments below half an ulp near
 * -86 mV); FP32 for the other models' FP32 path. */
int fw_u_is_f64(int32_t model, int32_t precision);

/* Weight builder: w[N*K] float64 raw weights (before the D*dt/dr**2 scale),
 * kind FW_STENCIL_ISO / ANISO / SYM; m int8 [N] with 0/1 values;
 * d: device pointers to the full half-step fields in the order above (sym:
 * its four corner-averaged arrays xx, xy, yx, yy).  w must be zeroed. */
int fw_compute_weights(int32_t ndim, int32_t kind, double* w, const int8_t* m,
                       const double* const* d, const int64_t* shape,
                       void* stream);

/* One diffusion sweep over the active indexes with precomputed weights:
 * u_new[c] = sum_k u[c + o_k] * w[c, k] (K = 5, 9, 7 or 19). */
int fw_diffusion(int32_t ndim, int32_t K, double* u_new, const double* u,
                 const double* w, const int64_t* indexes, int64_t n_indexes,
                 const int64_t* shape, void* stream);

/* One ionic sweep: state updated in place, u_new += dt * f(u, state)
 * (u_new -= dt * I_ion for the physiological models).  state: host array of
 * n_state device pointers in the reference kernel's argument order. */
int fw_ionic(int32_t model, int32_t precision, void* u_new, const void* u,
             void* const* state, int32_t n_state, const int64_t* indexes,
             int64_t n_indexes, double dt, const double* params,
             int32_t n_params, void* stream);

/* Fused stepping: plan creation validates the description and copies it. */
int fw_plan_create(const fw_plan_desc* desc, fw_plan** out);
int fw_plan_destroy(fw_plan* plan);

/* For FW_STENCIL_WEIGHTS plans: weights per node, 5 or 9 (2D), 7 or 19 (3D);
 * the default is 9 / 19. */
int fw_plan_set_weights_k(fw_plan* plan, int32_t K);

/* n_steps fused steps starting at global step `step0`: step s reads `u`,
 * writes `u_new`, then the two swap; after an odd n_steps the result is in
 * the buffer passed as u_new.  state: host array of device pointers. */
int fw_run(fw_plan* plan, void* u, void* u_new, void* const* state,
           int64_t step0, int32_t n_steps, const fw_track_desc* track,
           void* stream);

/* Same step as fw_run with n_steps = 1, recorded into a HIP graph of
 * n_steps nodes (even n_steps) and replayed; used for long stretches with
 * no host events.  Falls back to plain launches when capture fails. */
int fw_run_graph(fw_plan* plan, void* u, void* u_new, void* const* state,
                 int64_t step0, int32_t n_steps, const fw_track_desc* track,
                 void* stream);

/* Captures the graphs fw_run_graph would replay for the same arguments
 * (buffers, tracker, the plan's stimulus schedule, n_steps) and launches
 * nothing: capture and instantiation, which take milliseconds, then happen
 * outside a caller's timed or latency-bound region.  No counterpart in the
 * reference, which has no graphs. */
int fw_run_graph_prepare(fw_plan* plan, void* u, void* u_new, void* const* state,
                         int32_t n_steps, const fw_track_desc* track);

/* Same as fw_plan_create + fw_run + fw_plan_destroy with no stimulus
 * schedule, calling no hipMalloc, hipFree or other allocating routine: for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation.  No
 * counterpart in the reference.  The float copies and along-k tables of the
 * description's half-step fields (what fw_plan_create allocates) are staged
 * in a static device array of FW_NOALLOC_SCRATCH_BYTES on each GPU, kept
 * until the process ends; a description whose tables need more is refused
 * with FW_E_INVALID before any device work.  Every call on a GPU shares that
 * GPU's array and rewrites it before its steps read it: calls that overlap
 * on one GPU (different streams or threads) race and must be serialised by
 * the caller.  A FW_STENCIL_WEIGHTS description uses 9 (2D) / 19 (3D)
 * weights per node. */
#define FW_NOALLOC_SCRATCH_BYTES (4u << 20)
int fw_run_noalloc(const fw_plan_desc* desc, void* u, void* u_new, void* const* state,
                   int64_t step0, int32_t n_steps, const fw_track_desc* track,
                   void* stream);

int fw_stim_apply(fw_plan* plan, void* u, const fw_stim_desc* stim,
                  void* stream);

/* ---- stimulus schedule ------------------------------------------------------
 * StimSequence.stimulate_next (core/stimulation/stim_sequence.py:65-77) with
 * Stim.update_status (core/stimulation/stim.py:52-56) resolved ahead of time:
 * a stimulus with t >= stim.t that has not passed is applied at the start of
 * a step, then passes once t >= stim.t + duration, so with the FP64 time
 * fold t_s it is applied on the steps [step_lo, step_hi] (global step
 * numbers).  Applied to the potential before the step, in schedule order, to
 * nodes with FW_CODE_CONDUCT inside the box and the mask.  values: device
 * double [step_hi - step_lo + 1] or NULL; for a voltage stimulus, the value
 * of each step (StimVoltageListMatrix3D.stimulate,
 * cpuwave3D/stimulation/stim_voltage_list_matrix_3d.py:60-72). */
#define FW_MAX_STIMS 8
typedef struct fw_stim_sched {
  fw_stim_desc stim;
  int64_t step_lo, step_hi;
  const double* values;
} fw_stim_sched;

/* Replaces the plan's schedule with n (<= FW_MAX_STIMS) entries, copied;
 * masks and value lists stay the caller's device memory.  Graphs captured
 * with another schedule are dropped. */
int fw_plan_set_stims(fw_plan* plan, const fw_stim_sched* stims, int32_t n);

/* Applies the scheduled stimuli due at `step` to u (one launch each): the
 * first step of an fw_run / fw_run_graph call, whose kernels apply those of
 * the later steps. */
int fw_stim_apply_step(fw_plan* plan, void* u, int64_t step, void* stream);

/* One activation-time pass over all N nodes (act = step where u > thr first) */
int fw_act_time_update(fw_plan* plan, int32_t* act, const void* u, double thr,
                       int32_t step, void* stream);

/* ---- device trackers: read the potential as it stands before a step ------ */

/* LocalActivationTime2D/3DTracker._track and cross_threshold
 * (cpuwave2D/tracker/local_activation_time_2d_tracker.py:60-102): a node
 * crosses when u >= thr while not activated (its crossing step is appended to
 * steps[c * capacity + k] and it becomes activated) and is deactivated when
 * u < thr.  state: int32 [N], bit 0 activated, bits 1.. crossings so far,
 * zero-initialised by the caller.  *overflow (device) is set to 1 when a
 * node crosses more than `capacity` times.  The reference's layer structure
 * follows from the crossing steps and is rebuilt by the caller. */
int fw_lat_update(fw_plan* plan, const void* u, int32_t* state, int32_t* steps,
                  int32_t capacity, double thr, int32_t step, int32_t* overflow,
                  void* stream);

/* SpiralWaveCore2D/3DTracker._track (spiral_wave_core_2d_tracker.py:70-82,
 * 211-249; spiral_wave_core_3d_tracker.py:18-31): tips between u_prev and u
 * on every slice k, written as tips[5 * q + {0..4}] = (i, j, k, x, y) with
 * x = j + dy, y = i + dx as the reference appends them; *count (device) = number
 * found (entries beyond capacity are dropped; order is not the reference's,
 * sort by (k, i, j)). */
int fw_spiral_tips(fw_plan* plan, const void* u_prev, const void* u, double thr,
                   double* tips, int32_t capacity, int32_t* count, void* stream);

/* ECG2D/3DTracker.calc_ecg (ecg_2d_tracker.py:61-151, ecg_3d_tracker.py:
 * 100-140): out[e] = sum over updated nodes of (u_tr - u) / (d_e * dr), u_tr
 * the plan's diffusion of u, d_e the squared index distance to coords[3e..3e+2]
 * (2D: plus z^2).  coords, out: device.  Summation order differs from the
 * reference's parallel loop (block sums, FP64 atomics). */
int fw_ecg(fw_plan* plan, const void* u, const double* coords, int32_t n_points,
           double dr, double* out, void* stream);

/* ---- tissue set-up ---------------------------------------------------------
 * Fibrosis generated on the device, the FibrosisPattern classes' work
 * (core/fibrosis/fibrosis_pattern.py:17-51): writes mesh code 1 (tissue) or 2
 * (fibrosis) into the region [lo, hi) of an int8 mesh with `ndim` axes
 * `shape` (C order, device memory) and leaves the nodes outside it as they
 * are.  block null: each node is fibrosis with probability `density`
 * (Diffuse2DPattern / Diffuse3DPattern, cpuwave2D/fibrosis/
 * diffuse_2d_pattern.py:86-87); otherwise blocks of block[a] nodes along each
 * axis, tiling the region from lo and clipped at hi, are fibrosis with
 * probability `density` (Structural2DPattern / Structural3DPattern,
 * cpuwave2D/fibrosis/structural_2d_pattern.py:97-116).  The draws come from a
 * counter-based generator keyed by `seed` and the node's flat index (or the
 * block's index), so the mesh does not depend on the launch; they are not
 * numpy's or Python's streams, which only the host patterns reproduce. */
int fw_fibrosis(int8_t* mesh, int32_t ndim, const int64_t* shape, const int64_t* lo,
                const int64_t* hi, const int64_t* block, double density, uint64_t seed,
                void* stream);

/* ---- batches: many 2D tissues of one model, one launch per step ----------
 * No counterpart in the reference, which runs one tissue per process; this
 * is the ensemble / serving path built on the same fused step.  Each tissue
 * owns its device buffers and is described by one fw_tissue_desc; a block
 * table maps the launch's blocks to tissues; fw_batch_create and
 * fw_batch_set_active allocate device memory.  Local step of a
 * tissue = global step - step0; its u is in `u` at even local steps and in
 * `un` at odd ones.  act records local steps, probe[ls] = u at probe_idx
 * before local step ls, in the batch's precision (double or float). */
#define FW_BATCH_ALT_STATE 4
typedef struct fw_tissue_desc {
  void* u;
  void* un;
  void* state[FW_MAX_STATE];
  const uint8_t* code;
  int32_t* act;
  void* probe;
  int64_t ni, nj, n;
  double d[4]; /* uniform half-step fields: iso xx, yy; aniso xx, xy, yx, yy */
  double act_thr;
  int32_t stencil; /* FW_STENCIL_ISO or FW_STENCIL_ANISO */
  int32_t step0;
  int32_t act_start; /* first local step recorded in act, < 0 never */
  int32_t probe_idx; /* local flat node, < 0 none */
  int32_t probe_len; /* entries of probe; local steps beyond it are not sampled */
  int32_t reserved;
  /* models 0-4: a second set of the state arrays (same types and sizes),
   * which the multi-step launch shapes write between their two launches of
   * a pair; null entries make those shapes unavailable while the tissue is
   * active.  Their contents need no initialisation. */
  void* state_alt[FW_BATCH_ALT_STATE];
} fw_tissue_desc;

typedef struct fw_batch fw_batch;

/* Allocates the descriptor table for max_tissues slots. */
int fw_batch_create(int32_t model, int32_t precision, double D_model, double dt,
                    double dr2, const double* params, int32_t n_params,
                    int32_t max_tissues, fw_batch** out);
/* Frees the batch; no launch of it may still be pending (synchronize the
 * stream its calls used first). */
int fw_batch_destroy(fw_batch* batch);
/* Writes slot's descriptor (host struct copied to the device table).  For
 * models 5-7 it also builds the slot's node list (its updated nodes and its
 * probe node, int32, ascending) from the codes, in device memory the batch
 * owns and reuses; the slot must not be in use by a pending launch. */
int fw_batch_set_tissue(fw_batch* batch, int32_t slot, const fw_tissue_desc* desc,
                        void* stream);
/* Sets the tissues that fw_batch_run steps: `slots` (host, n_slots entries)
 * lists occupied slots.  The library builds the block table from their sizes
 * (a block covers fw_batch_block_nodes() consecutive nodes of one tissue)
 * and grows its device buffer when needed. */
int fw_batch_set_active(fw_batch* batch, const int32_t* slots, int32_t n_slots,
                        void* stream);
/* Empties a slot: it can no longer be listed in fw_batch_set_active or
 * stimulated until fw_batch_set_tissue fills it again. */
int fw_batch_release(fw_batch* batch, int32_t slot);
/* Nodes per block of the batch's current launch shape. */
int fw_batch_block_nodes(const fw_batch* batch);
/* Launch-shape variants of the batched step for a model (all give the same
 * results bit for bit): bit 0 = 4 consecutive nodes per work-item (models
 * 0-4), bit 1 = XCD-contiguous block map, bit 2 = state and new potential
 * accessed with the non-temporal cache hint.  Bit 3 (models 0-4) = K steps
 * per launch, K = 2 << bits 4-5, each wavefront marching down 64 columns by
 * 64 << bits 6-7 rows of a tissue; steps are run as pairs of such launches
 * (2K steps), the rest with the shape of bits 0-2.  Bit 8 (models 5-7) =
 * the shape of bits 1-2 over each tissue's node list instead of all its nodes
 * (no lane idles on a non-updated node).  Writes up to `capacity` of them to
 * `variants` and returns how many there are. */
int fw_batch_variants(int32_t model, int32_t* variants, int32_t capacity);
/* Selects the launch shape of fw_batch_run (rebuilds the block table when the
 * nodes per work-item or the node lists change); returns FW_E_INVALID for a
 * variant the model does not have. */
int fw_batch_set_variant(fw_batch* batch, int32_t variant, void* stream);
int fw_batch_get_variant(const fw_batch* batch);
/* n_steps global steps from step0; use_graph replays captured graphs
 * (recaptured when the block table or launch shape changes).  No stimulus
 * or admission may fall inside the run: the host applies them between
 * runs. */
int fw_batch_run(fw_batch* batch, int64_t step0, int32_t n_steps, int32_t use_graph,
                 void* stream);
/* Divergence check: flags[slot] (host, max_tissues entries) = 1 for every
 * active tissue whose potential before global step `step` holds a NaN or an
 * infinity, 0 otherwise.  One launch over the block table, reading each
 * tissue's current potential once; the copy to `flags` is enqueued on
 * `stream` (asynchronous for pinned host memory): read flags after the
 * stream has reached it. */
int fw_batch_nonfinite(fw_batch* batch, int64_t step, int32_t* flags, void* stream);
/* Applies a stimulus to slot's potential as it stands before global step
 * `step` (box in the tissue's own coordinates, third axis ignored). */
int fw_batch_stim(fw_batch* batch, int32_t slot, int64_t step, const fw_stim_desc* stim,
                  void* stream);

#ifdef __cplusplus
}
#endif
#endif
