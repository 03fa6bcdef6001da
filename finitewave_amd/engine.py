"""Device side of a simulation: HBM-resident state, the stepping plan and the
calls into ``libfinitewave_amd.so``.

State lives in torch tensors on the current HIP device (torch is only the
allocator and stream provider); every computation is a C-ABI call into the
hand-written gfx950 kernels.  Host numpy arrays on the model are the
reference's view of the state outside ``run()`` and on steps where a host
hook (command, saver, host tracker) needs them.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_int64, c_void_p

import numpy as np

from . import _native
from ._native import check, lib
from .stencil import compress_field


def _ptr(t) -> int:
    return t.data_ptr() if t is not None else 0


def _stream() -> int:
    import torch
    return torch.cuda.current_stream().cuda_stream


class Plan:
    """Owns one fw_plan and the device arrays it points to."""

    def __init__(self, ndim, model_id, stencil_kind, precision, shape, code_dev, fields,
                 weights_dev, K, D_model, dt, dr2, params):
        desc = _native.fw_plan_desc()
        desc.ndim = ndim
        desc.model = model_id
        desc.stencil = stencil_kind
        desc.precision = precision
        for a in range(3):
            desc.shape[a] = shape[a] if a < len(shape) else 1
        desc.code = _ptr(code_dev)
        desc.n_fields = len(fields)
        self._keep = [code_dev, weights_dev]
        for q, (mode, axis, value, data) in enumerate(fields):
            desc.fields[q].mode = mode
            desc.fields[q].axis = axis
            desc.fields[q].value = value
            desc.fields[q].data = _ptr(data)
            self._keep.append(data)
        desc.weights = _ptr(weights_dev)
        desc.D_model = float(D_model)
        desc.dt = float(dt)
        desc.dr2 = float(dr2)
        desc.n_params = len(params)
        for q, p in enumerate(params):
            desc.params[q] = float(p)
        self.desc = desc
        self.handle = c_void_p()
        check(lib().fw_plan_create(byref(desc), byref(self.handle)), "fw_plan_create")
        if stencil_kind == _native.STENCIL_WEIGHTS:
            check(lib().fw_plan_set_weights_k(self.handle, K), "fw_plan_set_weights_k")

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            lib().fw_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


def _upload_fields(torch, dev, arrays):
    fields = []
    for arr in arrays:
        mode, axis, value, data = compress_field(arr)
        dt = None if data is None else torch.from_numpy(np.ascontiguousarray(data)).to(dev)
        fields.append((mode, axis, value, dt))
    return fields


def raw_weights(stencil, model, tissue):
    """(*shape, K) float64 raw weights of an ISO / ANISO / SYM stencil, built
    on the GPU by fw_compute_weights, as a device tensor."""
    torch = _native.require_gpu()
    dev = torch.device("cuda", torch.cuda.current_device())
    mesh = np.asarray(tissue.mesh)
    shape = mesh.shape
    m01 = torch.from_numpy((mesh == 1).astype(np.int8)).to(dev)
    arrays = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
              for a in stencil.half_step_fields(model, tissue)]
    w = torch.zeros((*shape, stencil.K), dtype=torch.float64, device=dev)
    ptrs = (c_void_p * 9)(*[a.data_ptr() for a in arrays], *([0] * (9 - len(arrays))))
    shp = (c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))
    check(lib().fw_compute_weights(len(shape), stencil.kind, w.data_ptr(), m01.data_ptr(),
                                   ptrs, shp, _stream()), "fw_compute_weights")
    return w


def device_weights(stencil, model, tissue) -> np.ndarray:
    """The reference's ``compute_weights`` result (scaled, centre + 1)."""
    w = raw_weights(stencil, model, tissue)
    if stencil.kind == _native.STENCIL_SYM:
        # symmetric_stencil_2d.py:64-65
        w *= model.D_model * model.dt / model.dr ** 2
    else:
        # isotropic_stencil_2d.py:65-66, asymmetric_stencil_2d.py:81-82
        w = w * model.D_model * model.dt / model.dr ** 2
    w[..., stencil.centre] += 1
    return w.cpu().numpy()


def diffusion_kernel_for(stencil):
    """A callable with the reference's diffusion-kernel signature
    ``(u_new, u, w, indexes)`` running fw_diffusion; numpy arguments are
    copied to the device and the result written back into ``u_new``."""
    K = stencil.K

    def diffusion_kernel(u_new, u, w, indexes):
        torch = _native.require_gpu()
        dev = torch.device("cuda", torch.cuda.current_device())
        host = isinstance(u_new, np.ndarray)
        if host:
            t_un = torch.from_numpy(np.ascontiguousarray(u_new, dtype=np.float64)).to(dev)
            t_u = torch.from_numpy(np.ascontiguousarray(u, dtype=np.float64)).to(dev)
            t_w = torch.from_numpy(np.ascontiguousarray(w, dtype=np.float64)).to(dev)
            t_i = torch.from_numpy(np.ascontiguousarray(indexes, dtype=np.int64)).to(dev)
        else:
            t_un, t_u, t_w, t_i = u_new, u, w, indexes
        shape = tuple(t_u.shape)
        shp = (c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))
        check(lib().fw_diffusion(len(shape), K, t_un.data_ptr(), t_u.data_ptr(),
                                 t_w.data_ptr(), t_i.data_ptr(), int(t_i.numel()), shp,
                                 _stream()), "fw_diffusion")
        if host:
            u_new[...] = t_un.cpu().numpy()
        return u_new

    return diffusion_kernel


class Engine:
    """Device state and plan of one CardiacModel."""

    def __init__(self, model):
        self.torch = _native.require_gpu()
        self.model = model
        self.dev = self.torch.device("cuda", self.torch.cuda.current_device())
        self.plan = None
        self.swapped = False
        self.act = {}  # tracker -> int32 device tensor
        self.stim_sched = []  # fw_stim_sched entries of the current run
        self._stim_keep = []
        self.build()

    # -- geometry / plan ---------------------------------------------------
    def build(self) -> None:
        torch, model = self.torch, self.model
        tissue = model.cardiac_tissue
        tissue.check_boundary()
        mesh = np.asarray(tissue.mesh)
        self.shape = mesh.shape
        self.ndim = mesh.ndim
        self.N = int(mesh.size)
        self.precision = model._precision_id()
        # dtype of u / u_new (FP64 also on the FP32 path of LR91/TP06/CRN)
        self.rdtype = (torch.float64 if _native.u_is_f64(model._model_id(), self.precision)
                       else torch.float32)
        self.code_base = tissue.node_codes()
        self.code = torch.from_numpy(self.code_base).to(self.dev)
        if getattr(self, "probe_cells", None):
            self.set_probe_bits(self.probe_cells)
        stencil = model.stencil
        fields, weights, K = [], None, 0
        kind = getattr(stencil, "kind", None)
        if kind in (_native.STENCIL_ISO, _native.STENCIL_ANISO):
            fields = _upload_fields(torch, self.dev, stencil.half_step_fields(model, tissue))
            plan_kind = kind
        else:
            if kind == _native.STENCIL_SYM:
                w = torch.from_numpy(device_weights(stencil, model, tissue)).to(self.dev)
            else:  # any user stencil returning (*shape, K) weights
                w = torch.from_numpy(np.ascontiguousarray(
                    stencil.compute_weights(model, tissue), dtype=np.float64)).to(self.dev)
            K = int(w.shape[-1])
            weights = w.contiguous()
            plan_kind = _native.STENCIL_WEIGHTS
        self.model_id = model._model_id()
        params = model._param_values()
        if self.plan is not None:
            self.plan.close()
        self.plan = Plan(self.ndim, self.model_id, plan_kind, self.precision, self.shape,
                         self.code, fields, weights, K, model.D_model, model.dt,
                         model.dr ** 2, params)
        self.state_names = model._state_call_order()
        self.state_f64 = [_native.state_is_f64(self.model_id, self.precision, k)
                          for k in range(len(self.state_names))]
        self.state_off = [_native.state_offset(self.model_id, self.precision, k)
                          for k in range(len(self.state_names))]
        if self.stim_sched:
            self._push_stims()

    # -- stimulus schedule -------------------------------------------------------
    def set_stim_schedule(self, entries) -> None:
        """entries: [(fw_stim_desc, device value list or None, first step,
        last step)] in sequence order (fw_plan_set_stims)."""
        if len(entries) > _native.MAX_STIMS:
            raise ValueError(f"at most {_native.MAX_STIMS} scheduled stimuli")
        sched = []
        keep = []
        for desc, vals, lo, hi in entries:
            e = _native.fw_stim_sched()
            e.stim = desc
            e.step_lo = int(lo)
            e.step_hi = int(hi)
            e.values = vals.data_ptr() if vals is not None else 0
            sched.append(e)
            keep.append(vals)
        self.stim_sched = sched
        self._stim_keep = keep
        self._push_stims()

    def clear_stim_schedule(self) -> None:
        self.stim_sched = []
        self._stim_keep = []
        self._push_stims()

    def _push_stims(self) -> None:
        n = len(self.stim_sched)
        arr = (_native.fw_stim_sched * max(1, n))(*self.stim_sched)
        check(lib().fw_plan_set_stims(self.plan.handle, arr, n), "fw_plan_set_stims")

    def apply_stims_at(self, step: int) -> None:
        """The scheduled stimuli of `step` on the current potential."""
        if self.stim_sched:
            check(lib().fw_stim_apply_step(self.plan.handle, self.u_dev().data_ptr(),
                                           int(step), _stream()), "fw_stim_apply_step")

    def set_probe_bits(self, flat_cells) -> None:
        self.probe_cells = list(flat_cells)
        code = self.code_base.copy().reshape(-1)
        if len(flat_cells):
            code[np.asarray(flat_cells, dtype=np.int64)] |= _native.CODE_PROBE
        self.code.copy_(self.torch.from_numpy(code.reshape(self.shape)))

    # -- state transfer ------------------------------------------------------
    def upload(self) -> None:
        torch, m = self.torch, self.model
        self.A = torch.from_numpy(np.ascontiguousarray(m.u)).to(self.dev, self.rdtype)
        self.B = torch.from_numpy(np.ascontiguousarray(m.u_new)).to(self.dev, self.rdtype)
        self.swapped = False
        self.state = []
        for name, f64, off in zip(self.state_names, self.state_f64, self.state_off):
            dt = torch.float64 if f64 else torch.float32
            a = getattr(m, name)
            if off:  # stored as the difference from `off`, formed in double
                a = np.asarray(a, dtype=np.float64) - off
            self.state.append(torch.from_numpy(
                np.ascontiguousarray(a)).to(self.dev, dt).contiguous())
        self._state_ptrs = (c_void_p * max(1, len(self.state)))(
            *[s.data_ptr() for s in self.state])

    def u_dev(self):
        return self.B if self.swapped else self.A

    def u_new_dev(self):
        return self.A if self.swapped else self.B

    def state_values(self):
        """(name, host array) of every state array as the model holds it: the
        value, i.e. the stored difference plus its offset where there is one."""
        for name, t, off in zip(self.state_names, self.state, self.state_off):
            a = t.cpu().numpy()
            yield name, (a.astype(np.float64) + off if off else a)

    def download(self) -> None:
        m = self.model
        m.u = self.u_dev().cpu().numpy()
        m.u_new = self.u_new_dev().cpu().numpy()
        for name, a in self.state_values():
            setattr(m, name, a)

    # -- stepping ------------------------------------------------------------
    def run_steps(self, step0: int, n: int, act_tracker=None, act_flags=None,
                  probes=None, probe_rows=None, use_graph: bool = True,
                  stim_first: bool = True) -> None:
        """n fused steps starting at global step step0.  The scheduled
        stimuli of step0 are applied first unless stim_first is False (the
        caller applied them); those of the later steps are applied by the
        step kernels."""
        if n <= 0:
            return
        L = lib()
        track = _native.fw_track_desc()
        keep = []
        if self.stim_sched:
            track.stim_end = step0 + n
        if act_tracker is not None:
            act = self.act[act_tracker]
            track.act = act.data_ptr()
            track.act_thr = float(act_tracker.threshold)
            flags = np.ascontiguousarray(act_flags, dtype=np.uint8)
            keep.append(flags)
            track.act_flags = flags.ctypes.data
        if probes is not None:
            rows = np.ascontiguousarray(probe_rows, dtype=np.int32)
            keep.append(rows)
            track.n_probe = probes.n
            track.probe_idx = probes.idx.data_ptr()
            track.probe_var = probes.var.data_ptr()
            track.probe_buf = probes.buf.data_ptr()
            track.probe_rows = rows.ctypes.data
        stream = _stream()
        uniform = act_flags is None or (np.all(act_flags == act_flags[0]))
        graph_ok = use_graph and probes is None and uniform and n >= 2
        u_ptr, un_ptr = self.u_dev().data_ptr(), self.u_new_dev().data_ptr()
        if self.stim_sched and stim_first:
            # launched last before the steps: no host work between the two
            check(L.fw_stim_apply_step(self.plan.handle, u_ptr, int(step0), stream),
                  "fw_stim_apply_step")
        if graph_ok:
            # graphs exist for both buffer orientations; an odd n ends with
            # one plain launch inside fw_run_graph
            check(L.fw_run_graph(self.plan.handle, u_ptr, un_ptr, self._state_ptrs, step0, n,
                                 byref(track), stream), "fw_run_graph")
            if n % 2:
                self.swapped = not self.swapped
            return
        self._plain(track, step0, n, act_flags, 0, stream)

    def prepare_graphs(self, n: int, act_tracker=None, act_on: bool = True) -> None:
        """Captures the graphs a graph-replayed run_steps of n steps from the
        current buffer orientation would use (fw_run_graph_prepare),
        launching nothing; run_steps then only replays them."""
        pairs = n // 2
        if pairs <= 0:
            return
        track = _native.fw_track_desc()
        track.stim_end = 0
        flags = np.full(2 * pairs, 1 if act_on else 0, dtype=np.uint8)
        if act_tracker is not None:
            track.act = self.act[act_tracker].data_ptr()
            track.act_thr = float(act_tracker.threshold)
            track.act_flags = flags.ctypes.data
        check(lib().fw_run_graph_prepare(self.plan.handle, self.u_dev().data_ptr(),
                                         self.u_new_dev().data_ptr(), self._state_ptrs,
                                         2 * pairs, byref(track)), "fw_run_graph_prepare")

    def _plain(self, track, step0, n, act_flags, offset, stream) -> None:
        sub = _native.fw_track_desc.from_buffer_copy(track)
        if track.act_flags:
            sub.act_flags = track.act_flags + offset
        if track.probe_rows:
            sub.probe_rows = track.probe_rows + 4 * offset
        u, un = self.u_dev(), self.u_new_dev()
        check(lib().fw_run(self.plan.handle, u.data_ptr(), un.data_ptr(), self._state_ptrs,
                           step0, n, byref(sub), stream), "fw_run")
        if n % 2:
            self.swapped = not self.swapped

    def apply_stim(self, desc) -> None:
        check(lib().fw_stim_apply(self.plan.handle, self.u_dev().data_ptr(), byref(desc),
                                  _stream()), "fw_stim_apply")

    def act_update(self, tracker, step: int) -> None:
        check(lib().fw_act_time_update(self.plan.handle, self.act[tracker].data_ptr(),
                                       self.u_dev().data_ptr(), float(tracker.threshold),
                                       int(step), _stream()), "fw_act_time_update")

    def close(self) -> None:
        if self.plan is not None:
            self.plan.close()
            self.plan = None

    def __deepcopy__(self, memo):
        return None


class Probes:
    """Device-side probe list shared by the point-probe trackers."""

    def __init__(self, engine, entries, rows_capacity: int):
        torch = engine.torch
        self.entries = entries  # list of (flat cell, var slot)
        self.n = len(entries)
        self.idx = torch.tensor([e[0] for e in entries], dtype=torch.int64, device=engine.dev)
        self.var = torch.tensor([e[1] for e in entries], dtype=torch.int32, device=engine.dev)
        self.capacity = rows_capacity
        self.buf = torch.zeros((rows_capacity, max(1, self.n)), dtype=torch.float64,
                               device=engine.dev)
