"""ctypes binding of ``libfinitewave_amd.so`` (declared in include/finitewave_amd.h).

The library is the only compute path: if it is missing or a GPU is not
available every stepping entry point raises ``NativeUnavailableError``; there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_int32, c_int64, c_void_p, POINTER
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libfinitewave_amd.so"

FW_OK = 0
MODEL_IDS = {
    "aliev_panfilov": 0, "barkley": 1, "mitchell_schaeffer": 2, "fenton_karma": 3,
    "bueno_orovio": 4, "luo_rudy91": 5, "tp06": 6, "courtemanche": 7,
}
STENCIL_ISO, STENCIL_ANISO, STENCIL_WEIGHTS, STENCIL_SYM = 0, 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1
FIELD_UNIFORM, FIELD_AXIS, FIELD_FULL = 0, 1, 2
CODE_CONDUCT, CODE_FROZEN, CODE_PROBE = 1, 2, 4
MAX_STATE, MAX_PARAMS, MAX_FIELDS, MAX_STIMS = 24, 64, 9, 8
NOALLOC_SCRATCH_BYTES = 4 << 20
BATCH_ALT_STATE = 4
ABI_VERSION = 4

# every symbol include/finitewave_amd.h declares
EXPORTS = (
    "fw_abi_version", "fw_last_error", "fw_model_info", "fw_state_is_f64", "fw_state_offset",
    "fw_u_is_f64",
    "fw_compute_weights", "fw_diffusion", "fw_ionic", "fw_plan_create",
    "fw_plan_destroy", "fw_plan_set_weights_k", "fw_run", "fw_run_graph",
    "fw_run_graph_prepare", "fw_run_noalloc",
    "fw_stim_apply", "fw_plan_set_stims", "fw_stim_apply_step", "fw_act_time_update",
    "fw_lat_update", "fw_spiral_tips", "fw_ecg",
    "fw_fibrosis",
    "fw_batch_create", "fw_batch_destroy",
    "fw_batch_set_tissue", "fw_batch_set_active", "fw_batch_release", "fw_batch_block_nodes",
    "fw_batch_variants", "fw_batch_set_variant", "fw_batch_get_variant",
    "fw_batch_run",
    "fw_batch_stim", "fw_batch_nonfinite",
)


class NativeUnavailableError(RuntimeError):
    """The HIP library is not built or cannot be loaded."""


class FwError(RuntimeError):
    """A C-ABI call returned a negative status."""


class fw_field(ctypes.Structure):
    _fields_ = [("mode", c_int32), ("axis", c_int32), ("value", c_double), ("data", c_void_p)]


class fw_plan_desc(ctypes.Structure):
    _fields_ = [("ndim", c_int32), ("model", c_int32), ("stencil", c_int32),
                ("precision", c_int32), ("shape", c_int64 * 3), ("code", c_void_p),
                ("n_fields", c_int32), ("fields", fw_field * MAX_FIELDS),
                ("weights", c_void_p), ("D_model", c_double), ("dt", c_double),
                ("dr2", c_double), ("n_params", c_int32), ("params", c_double * MAX_PARAMS)]


class fw_track_desc(ctypes.Structure):
    _fields_ = [("act", c_void_p), ("act_thr", c_double), ("act_flags", c_void_p),
                ("n_probe", c_int32), ("probe_idx", c_void_p), ("probe_var", c_void_p),
                ("probe_buf", c_void_p), ("probe_rows", c_void_p), ("stim_end", c_int64)]


class fw_stim_desc(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("lo", c_int64 * 3), ("hi", c_int64 * 3),
                ("mask", c_void_p), ("value", c_double), ("clamp", c_int32),
                ("u_max", c_double)]


class fw_stim_sched(ctypes.Structure):
    _fields_ = [("stim", fw_stim_desc), ("step_lo", c_int64), ("step_hi", c_int64),
                ("values", c_void_p)]


_lib = None


def lib():
    """Load the library once; raise NativeUnavailableError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NativeUnavailableError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    try:
        L = ctypes.CDLL(str(LIB_PATH))
    except OSError as e:  # pragma: no cover - depends on the ROCm runtime
        raise NativeUnavailableError(f"cannot load {LIB_PATH}: {e}") from e
    vp, i32, i64, dbl = c_void_p, c_int32, c_int64, c_double
    L.fw_abi_version.restype = ctypes.c_int
    L.fw_last_error.restype = ctypes.c_char_p
    L.fw_model_info.argtypes = [i32, POINTER(i32), POINTER(i32)]
    L.fw_state_is_f64.argtypes = [i32, i32, i32]
    L.fw_state_offset.argtypes = [i32, i32, i32, POINTER(dbl)]
    L.fw_u_is_f64.argtypes = [i32, i32]
    L.fw_compute_weights.argtypes = [i32, i32, vp, vp, POINTER(vp), POINTER(i64), vp]
    L.fw_diffusion.argtypes = [i32, i32, vp, vp, vp, vp, i64, POINTER(i64), vp]
    L.fw_ionic.argtypes = [i32, i32, vp, vp, POINTER(vp), i32, vp, i64, dbl, POINTER(dbl),
                           i32, vp]
    L.fw_plan_create.argtypes = [POINTER(fw_plan_desc), POINTER(vp)]
    L.fw_plan_destroy.argtypes = [vp]
    L.fw_plan_set_weights_k.argtypes = [vp, i32]
    L.fw_run.argtypes = [vp, vp, vp, POINTER(vp), i64, i32, POINTER(fw_track_desc), vp]
    L.fw_run_graph.argtypes = [vp, vp, vp, POINTER(vp), i64, i32, POINTER(fw_track_desc), vp]
    L.fw_run_graph_prepare.argtypes = [vp, vp, vp, POINTER(vp), i32, POINTER(fw_track_desc)]
    L.fw_run_noalloc.argtypes = [POINTER(fw_plan_desc), vp, vp, POINTER(vp), i64, i32,
                                 POINTER(fw_track_desc), vp]
    L.fw_stim_apply.argtypes = [vp, vp, POINTER(fw_stim_desc), vp]
    L.fw_plan_set_stims.argtypes = [vp, POINTER(fw_stim_sched), i32]
    L.fw_stim_apply_step.argtypes = [vp, vp, i64, vp]
    L.fw_act_time_update.argtypes = [vp, vp, vp, dbl, i32, vp]
    L.fw_lat_update.argtypes = [vp, vp, vp, vp, i32, dbl, i32, vp, vp]
    L.fw_spiral_tips.argtypes = [vp, vp, vp, dbl, vp, i32, vp, vp]
    L.fw_ecg.argtypes = [vp, vp, vp, i32, dbl, vp, vp]
    L.fw_fibrosis.argtypes = [vp, i32, POINTER(i64), POINTER(i64), POINTER(i64), vp, dbl,
                              ctypes.c_uint64, vp]
    for name in EXPORTS:
        if name not in ("fw_last_error",):
            getattr(L, name).restype = ctypes.c_int
    if L.fw_abi_version() != ABI_VERSION:
        raise NativeUnavailableError(
            f"{LIB_PATH} has ABI version {L.fw_abi_version()}, this package needs "
            f"{ABI_VERSION}; rebuild it")
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != FW_OK:
        msg = lib().fw_last_error().decode(errors="replace")
        raise FwError(f"{what or 'finitewave_amd'} failed ({rc}): {msg}")


def model_info(model_id: int) -> tuple[int, int]:
    ns, npar = c_int32(), c_int32()
    check(lib().fw_model_info(model_id, ctypes.byref(ns), ctypes.byref(npar)), "fw_model_info")
    return ns.value, npar.value


def state_is_f64(model_id: int, precision: int, k: int) -> bool:
    rc = lib().fw_state_is_f64(model_id, precision, k)
    if rc < 0:
        check(rc, "fw_state_is_f64")
    return bool(rc)


def state_offset(model_id: int, precision: int, k: int) -> float:
    """The value state k is stored relative to (the array holds value - offset)."""
    out = c_double()
    check(lib().fw_state_offset(model_id, precision, k, ctypes.byref(out)), "fw_state_offset")
    return out.value


def u_is_f64(model_id: int, precision: int) -> bool:
    """Whether u / u_new are stored as float64 for this model and precision."""
    rc = lib().fw_u_is_f64(model_id, precision)
    if rc < 0:
        check(rc, "fw_u_is_f64")
    return bool(rc)


def require_gpu():
    """torch with a visible HIP device, or NativeUnavailableError."""
    import torch
    if not torch.cuda.is_available():
        raise NativeUnavailableError("no HIP device is visible; the stepping engine runs only "
                                     "on the GPU")
    lib()
    return torch
