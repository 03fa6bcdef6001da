"""Batches: many 2D tissues of one ionic model stepped by one launch per step.

A ``TissueBatch`` owns a ``fw_batch`` (descriptor table + block table) and
the HBM buffers of every admitted tissue.  Tissues of uneven size, fibrosis
and stimulus protocol share each launch; between launches the host admits
new tissues, retires finished ones and applies stimuli.  Stretches of steps
without a stimulus are replayed from a captured two-step HIP graph.

Per tissue the semantics are those of ``CardiacModel.run()``: the model's
initial state and ``u_new`` initialisation, the reference's stimulus
fencepost (a stimulus applies on every step with ``t >= t0`` until the first
step with ``t >= t0 + duration``), model time as the left fold ``t += dt``,
``ceil(t_max/dt)`` steps.  The activation-time map records the first step
with ``u > threshold`` from ``act_start`` on; the probe records ``u`` at one
node before every step.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _native
from ._native import check, lib
from .analysis import apd90, captured, conduction_velocity, front_velocity


class _Shim:
    """What a Stim's device descriptor builder needs from an engine."""

    def __init__(self, torch, dev, shape):
        self.torch, self.dev, self.shape = torch, dev, shape


@dataclass
class BatchSim:
    """One tissue of a batch."""
    tissue: object                 # CardiacTissue2D (mesh, scalar conductivity, fibers)
    t_max: float
    stims: list = field(default_factory=list)
    probe: tuple | None = None     # (i, j) node sampled every step
    act_threshold: float | None = None
    act_start: float = 0.0         # model time from which activation is recorded
    apd_threshold: float | None = None
    apd_start: float = 0.0         # model time from which the trace is analysed (S2)
    cv_row: int | None = None      # row for the conduction-velocity estimate
    tag: object = None
    # filled by the batch
    slot: int = -1
    step0: int = 0
    n_steps: int = 0
    result: dict | None = None


def _uniform_fields(sim, stencil_kind, D_al=1, D_ac=1 / 9):
    """The four (iso: two) half-step values of a tissue whose conductivity is
    a scalar and whose fibres are one direction; the reference's arithmetic
    (asymmetric_stencil_2d.py:97-161) evaluated on one node."""
    cond = sim.tissue.conductivity
    if np.ndim(cond) != 0:
        raise ValueError("batched tissues need a scalar conductivity")
    c = float(cond)
    if stencil_kind == _native.STENCIL_ISO:
        d = 0.5 * (c + c)
        return [d, d, 0.0, 0.0]
    f = np.asarray(sim.tissue.fibers, dtype=np.float64).reshape(-1, 2)
    if not np.all(f == f[0]):
        raise ValueError("batched tissues need one fibre direction")
    out = []
    for axis in range(2):
        for comp in range(2):
            D = D_ac * (axis == comp) + (D_al - D_ac) * f[0, axis] * f[0, comp]
            out.append(0.5 * (D * c + D * c))
    return out


def time_fold(dt: float, n: int) -> np.ndarray:
    """Model time before each of n steps: the left fold t += dt from 0."""
    out = np.empty(n)
    t = 0.0
    for k in range(n):
        out[k] = t
        t += dt
    return out


def run_length(t_max: float, dt: float, fold: np.ndarray) -> int:
    """Steps model.run() takes to t_max: ceil(t_max/dt), or fewer when the
    folded time passes t_max first (cardiac_model.py:191-203)."""
    n = int(math.ceil(t_max / dt))
    over = np.flatnonzero(fold[1:n + 1] > t_max)
    return int(over[0]) + 1 if over.size else n


def stim_schedule(stims, n_steps: int, fold: np.ndarray) -> list:
    """(local step, stim) pairs in application order: a stimulus applies on
    every step with t >= t0 until and including the first step with
    t >= t0 + duration; stimuli due on one step apply in list order
    (stim_sequence.py:74-77, stim.py:52-56)."""
    ev = []
    for order, st in enumerate(stims):
        t0 = st.t
        dur = getattr(st, "duration", 0.0)
        first = int(np.searchsorted(fold[:n_steps], t0, side="left"))
        if first >= n_steps:
            continue
        last = first
        while last < n_steps - 1 and not (fold[last] >= t0 + dur):
            last += 1
        for k in range(first, last + 1):
            ev.append((k, order, st))
    ev.sort(key=lambda e: (e[0], e[1]))
    return [(k, st) for k, _, st in ev]


_XCD_BIT = 2  # launch-shape bit of the XCD-contiguous block map (fw_batch.hpp)
_TUNE_CACHE: dict = {}


def _tune_cache_path() -> Path:
    return Path(os.environ.get("FW_AUTOTUNE_CACHE",
                               Path.home() / ".cache" / "finitewave_amd" / "batch_tune.json"))


def _tune_cache_load() -> None:
    if _TUNE_CACHE:
        return
    try:
        _TUNE_CACHE.update(json.loads(_tune_cache_path().read_text()))
    except (OSError, ValueError):
        pass


def _tune_cache_store(key: str, variant: int, timings: dict) -> None:
    _TUNE_CACHE[key] = variant
    p = _tune_cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        data = {}
        if p.exists():
            data = json.loads(p.read_text())
        data[key] = variant
        data[key + "|ms"] = {str(v): t for v, t in timings.items()}
        p.write_text(json.dumps(data, indent=1, sort_keys=True))
    except (OSError, ValueError):
        pass  # the cache is an optimisation; a read-only home does not stop the run


def _on_stream(method):
    """Runs a TissueBatch method with the batch's own stream current, so that
    its allocations, copies and launches are ordered on that stream."""
    def wrapped(self, *args, **kwargs):
        with self.torch.cuda.stream(self.stream):
            return method(self, *args, **kwargs)
    wrapped.__name__, wrapped.__doc__ = method.__name__, method.__doc__
    return wrapped


class TissueBatch:
    """Many tissues of `model_cls` (a finitewave_amd 2D model class).

    Launch shape: the batched kernel has variants (nodes per work-item,
    block-to-XCD map; `fw_batch_variants`) that compute the same results bit
    for bit.  With ``autotune`` the batch times each on its own next steps
    (``tune_steps`` each after one untimed step), keeps the fastest and
    remembers it per (device, model, precision, size class) in memory and in
    a JSON cache (``FW_AUTOTUNE_CACHE``, default
    ``~/.cache/finitewave_amd/batch_tune.json``).  The size class is the
    power of two of the batch's node count; admissions and retirements that
    change it select the shape again (from the cache, or by timing), since
    the best shape of a few small tissues is not that of many."""

    def __init__(self, model_cls, precision: str = "fp32", dt: float = 0.01, dr: float = 0.25,
                 max_tissues: int = 1024, use_graph: bool = True, chunk_cap: int = 4000,
                 autotune: bool = True, tune_steps: int = 16):
        self.torch = _native.require_gpu()
        self.dev = self.torch.device("cuda", self.torch.cuda.current_device())
        # every batch works on a stream of its own (default priority): the
        # batches of a server then overlap on the device
        self.stream = self.torch.cuda.Stream(device=self.dev)
        self.proto = model_cls()
        self.proto.precision = precision
        self.prec = self.proto._precision_id()
        self.model_id = self.proto._model_id()
        self.dt, self.dr = dt, dr
        self.use_graph = use_graph
        self.chunk_cap = chunk_cap
        self.max_tissues = max_tissues
        params = self.proto._param_values()
        par = (ctypes.c_double * len(params))(*params)
        self.handle = ctypes.c_void_p()
        L = lib()
        L.fw_batch_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_int32,
                                      ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
        L.fw_batch_destroy.argtypes = [ctypes.c_void_p]
        L.fw_batch_set_tissue.argtypes = [ctypes.c_void_p, ctypes.c_int32,
                                          ctypes.POINTER(_TissueDesc), ctypes.c_void_p]
        L.fw_batch_set_active.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                          ctypes.c_void_p]
        L.fw_batch_release.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.fw_batch_variants.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32]
        L.fw_batch_set_variant.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        L.fw_batch_get_variant.argtypes = [ctypes.c_void_p]
        L.fw_batch_block_nodes.argtypes = [ctypes.c_void_p]
        L.fw_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_void_p]
        L.fw_batch_stim.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                    ctypes.POINTER(_native.fw_stim_desc), ctypes.c_void_p]
        L.fw_batch_nonfinite.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_void_p]
        check(L.fw_batch_create(self.model_id, self.prec, float(self.proto.D_model), dt,
                                dr ** 2, par, len(params), max_tissues,
                                ctypes.byref(self.handle)), "fw_batch_create")
        self.autotune, self.tune_steps = autotune, tune_steps
        self.tuned = not autotune
        self.tuning: dict | None = None   # variant -> ms per step, when measured here
        self._tuned_key: str | None = None  # size class the current shape was chosen for
        self.step = 0                     # global step of the next launch
        self.sims: dict[int, BatchSim] = {}
        self.buffers: dict[int, dict] = {}
        self.events: dict[int, list] = {}  # slot -> sorted [(local step, stim)]
        self.free = list(range(max_tissues - 1, -1, -1))
        self.blocks_dirty = True
        self.done: list[BatchSim] = []
        self._fold = np.zeros(1)
        self.names = self.proto._state_call_order()
        self.f64 = [_native.state_is_f64(self.model_id, self.prec, k)
                    for k in range(len(self.names))]
        self.off = [_native.state_offset(self.model_id, self.prec, k)
                    for k in range(len(self.names))]
        # the multi-step launch shapes (light models) write a second state set
        self.alt_state = self.model_id <= 4
        self._flags = None   # pinned host flags of the divergence check
        self._check = None   # (event, sims) of the check in flight

    # ------------------------------------------------------------------ time
    def _t_fold(self, n: int) -> np.ndarray:
        """Model time before each of the first n local steps (t += dt)."""
        if len(self._fold) < n + 1:
            self._fold = time_fold(self.dt, max(n + 1, 2 * len(self._fold)))
        return self._fold

    def footprint(self, shape, horizon_steps) -> int:
        """Device bytes a tissue of `shape` needs (state and, for the light
        models, its second set; codes, maps, probe)."""
        n = int(np.prod(shape))
        rb = 8 if _native.u_is_f64(self.model_id, self.prec) else 4
        st = sum(8 if f else 4 for f in self.f64)
        if self.alt_state:
            st *= 2
        # models 5-7: the library's list of updated nodes (int32, at most one per node)
        lst = 4 if self.model_id >= 5 else 0
        return n * (2 * rb + st + 1 + 4 + lst) + rb * (horizon_steps + 1)

    # ------------------------------------------------------------------ admission
    @_on_stream
    def admit(self, sim: BatchSim) -> int:
        if not self.free:
            raise RuntimeError("batch is full")
        torch = self.torch
        tissue = sim.tissue
        mesh = np.asarray(tissue.mesh)
        if mesh.ndim != 2:
            raise ValueError("batches hold 2D tissues")
        tissue.check_boundary()
        if tissue.special_boundaries is not None and np.any(tissue.special_boundaries):
            raise ValueError("batched tissues take no special boundaries")
        shape = mesh.shape
        n = int(mesh.size)
        slot = self.free.pop()
        stencil = _native.STENCIL_ISO if tissue.fibers is None else _native.STENCIL_ANISO
        d4 = _uniform_fields(sim, stencil)
        p = self.proto
        rd = torch.float64 if _native.u_is_f64(self.model_id, self.prec) else torch.float32
        u = torch.full(shape, float(p.init_u), dtype=rd, device=self.dev)
        un = u.clone() if p._copy_u_new else torch.zeros(shape, dtype=rd, device=self.dev)
        inits = {name: getattr(p, init_name) for name, init_name, _ in p._state}
        state = [torch.full(shape, float(inits[nm]) - off, device=self.dev,
                            dtype=torch.float64 if f else torch.float32)
                 for nm, f, off in zip(self.names, self.f64, self.off)]
        state_alt = [torch.empty_like(x) for x in state] if self.alt_state else []
        code = torch.from_numpy(tissue.node_codes()).to(self.dev)
        n_steps = run_length(sim.t_max, self.dt, self._t_fold(int(math.ceil(sim.t_max / self.dt)) + 1))
        fold = self._t_fold(n_steps + 1)
        act = None
        act_start = -1
        if sim.act_threshold is not None:
            act = torch.full(shape, -1, dtype=torch.int32, device=self.dev)
            act_start = int(np.searchsorted(fold[:n_steps], sim.act_start, side="left"))
        probe = None
        probe_idx = -1
        if sim.probe is not None:
            probe = torch.zeros(n_steps + 1, dtype=rd, device=self.dev)
            probe_idx = int(np.ravel_multi_index(tuple(sim.probe), shape))
        desc = _TissueDesc()
        desc.u, desc.un = u.data_ptr(), un.data_ptr()
        for k, s in enumerate(state):
            desc.state[k] = s.data_ptr()
        for k, s in enumerate(state_alt):
            desc.state_alt[k] = s.data_ptr()
        desc.code = code.data_ptr()
        desc.act = act.data_ptr() if act is not None else 0
        desc.probe = probe.data_ptr() if probe is not None else 0
        desc.ni, desc.nj, desc.n = shape[0], shape[1], n
        for q in range(4):
            desc.d[q] = d4[q]
        desc.stencil = stencil
        desc.step0 = self.step
        desc.act_start = act_start
        desc.probe_idx = probe_idx
        desc.act_thr = float(sim.act_threshold if sim.act_threshold is not None else 0.0)
        desc.probe_len = n_steps + 1 if probe is not None else 0
        check(lib().fw_batch_set_tissue(self.handle, slot, ctypes.byref(desc),
                                        self._stream()), "fw_batch_set_tissue")
        sim.slot, sim.step0, sim.n_steps = slot, self.step, n_steps
        self.sims[slot] = sim
        self.buffers[slot] = dict(u=u, un=un, state=state, state_alt=state_alt, code=code,
                                  act=act, probe=probe, shape=shape, desc=desc)
        self.events[slot] = stim_schedule(sim.stims, n_steps, fold)
        self.blocks_dirty = True
        return slot

    @_on_stream
    def retire(self, slot: int) -> BatchSim:
        sim = self.sims.pop(slot)
        buf = self.buffers.pop(slot)
        self.events.pop(slot, None)
        fold = self._t_fold(sim.n_steps + 1)
        res = {}
        ls_end = sim.n_steps
        u_final = buf["un"] if ls_end % 2 else buf["u"]
        res["u"] = u_final.cpu().numpy()
        if buf["act"] is not None:
            steps = buf["act"].cpu().numpy()
            at = -np.ones(steps.shape)
            hit = steps >= 0
            at[hit] = fold[steps[hit]]
            res["act_t"] = at
            if sim.cv_row is not None:
                res["cv"] = conduction_velocity(at, sim.cv_row, self.dr)
            fv = front_velocity(at, self.dr)
            res["cv_front"] = float(np.mean(fv)) if fv is not None else None
        if buf["probe"] is not None:
            trace = buf["probe"][:sim.n_steps].cpu().numpy().astype(np.float64)
            res["trace"] = trace
            if sim.apd_threshold is not None:
                # the S2 run's own trace (RestitutionCurve.ipynb cell 6): from
                # the first step with t >= apd_start on
                s0 = int(np.searchsorted(fold[:sim.n_steps], sim.apd_start, side="left"))
                res["captured"] = captured(trace, sim.apd_threshold, s0)
                res["apd90"] = apd90(trace, self.dt, sim.apd_threshold, s0)
        sim.result = res
        check(lib().fw_batch_release(self.handle, slot), "fw_batch_release")
        self.free.append(slot)
        self.blocks_dirty = True
        return sim

    @_on_stream
    def drop(self, slot: int) -> BatchSim:
        """Removes a tissue without reading its results (a cancelled or
        diverged request); its buffers are freed once the batch's stream has
        finished with them."""
        sim = self.sims.pop(slot)
        self.stream.synchronize()
        self.buffers.pop(slot)
        self.events.pop(slot, None)
        check(lib().fw_batch_release(self.handle, slot), "fw_batch_release")
        self.free.append(slot)
        self.blocks_dirty = True
        return sim

    @_on_stream
    def check_nonfinite(self) -> None:
        """Enqueues a divergence check of every tissue at the current step
        (fw_batch_nonfinite into pinned host memory); poll_nonfinite reads it
        once the stream has reached it."""
        if not self.sims or self._check is not None:
            return
        if self.blocks_dirty:
            self._rebuild_blocks()
        torch = self.torch
        if self._flags is None:
            self._flags = torch.zeros(self.max_tissues, dtype=torch.int32, pin_memory=True)
        check(lib().fw_batch_nonfinite(self.handle, self.step, self._flags.data_ptr(),
                                       self._stream()), "fw_batch_nonfinite")
        ev = torch.cuda.Event()
        ev.record(self.stream)
        self._check = (ev, dict(self.sims))

    def poll_nonfinite(self, wait: bool = False) -> list:
        """Slots found non-finite by the last check once it has completed
        (tissues that left the batch since are skipped); [] while it runs."""
        if self._check is None:
            return []
        ev, sims = self._check
        if not wait and not ev.query():
            return []
        ev.synchronize()
        self._check = None
        flags = self._flags.numpy()
        return [slot for slot, sim in sims.items()
                if flags[slot] and self.sims.get(slot) is sim]

    def nonfinite(self) -> list:
        """Slots whose potential holds a NaN or an infinity now."""
        self.check_nonfinite()
        return self.poll_nonfinite(wait=True)

    # ------------------------------------------------------------------ stepping
    def _stream(self):
        return self.stream.cuda_stream

    def _rebuild_blocks(self) -> None:
        slots = np.array(sorted(self.sims), dtype=np.int32)
        check(lib().fw_batch_set_active(self.handle, slots.ctypes.data, len(slots),
                                        self._stream()), "fw_batch_set_active")
        self.blocks_dirty = False

    def _apply_due_stims(self) -> None:
        for slot, sim in self.sims.items():
            ev = self.events[slot]
            ls = self.step - sim.step0
            while ev and ev[0][0] == ls:
                _, st = ev.pop(0)
                shim = _Shim(self.torch, self.dev, self.buffers[slot]["shape"])
                for d in st._descs(_ModelShim(self.dt), shim):
                    check(lib().fw_batch_stim(self.handle, slot, self.step, ctypes.byref(d),
                                              self._stream()), "fw_batch_stim")

    # ------------------------------------------------------------------ launch shape
    def variants(self) -> list:
        L = lib()
        n = L.fw_batch_variants(self.model_id, None, 0)
        arr = (ctypes.c_int32 * max(1, n))()
        L.fw_batch_variants(self.model_id, arr, n)
        return list(arr[:n])

    @property
    def variant(self) -> int:
        return lib().fw_batch_get_variant(self.handle)

    @_on_stream
    def set_variant(self, variant: int) -> None:
        check(lib().fw_batch_set_variant(self.handle, int(variant), self._stream()),
              "fw_batch_set_variant")

    def _tune_key(self) -> str:
        n = sum(int(self.buffers[s]["desc"].n) for s in self.sims)
        dev = self.torch.cuda.get_device_name(self.dev)
        return f"{dev}|m{self.model_id}|p{self.prec}|n2^{int(math.log2(max(1, n)))}"

    def _tune(self, n_avail: int) -> int:
        """Times every variant on the next steps; returns the steps run."""
        _tune_cache_load()
        key = self._tune_key()
        if key in _TUNE_CACHE:
            self.set_variant(_TUNE_CACHE[key])
            self.tuned, self._tuned_key = True, key
            return 0
        vs = self.variants()
        per = self.tune_steps + 1
        if len(vs) < 2:
            self.tuned, self._tuned_key = True, key
            return 0
        if n_avail < per * len(vs):
            return 0
        torch = self.torch
        L = lib()
        timings = {}
        step = self.step
        for v in vs:
            self.set_variant(v)
            check(L.fw_batch_run(self.handle, step, 1, 0, self._stream()), "fw_batch_run")
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            check(L.fw_batch_run(self.handle, step + 1, self.tune_steps, 0, self._stream()),
                  "fw_batch_run")
            e1.record()
            e1.synchronize()
            timings[v] = e0.elapsed_time(e1) / self.tune_steps
            step += per
        best = min(timings, key=timings.get)
        # among near-ties (within 3 %) prefer the XCD-contiguous block map: it
        # reads fewer bytes from HBM (neighbour rows stay in one XCD's L2),
        # which step time alone does not show when the Infinity Cache hides
        # the re-reads
        near = [v for v in vs if timings[v] <= 1.03 * timings[best] and v & _XCD_BIT]
        if near and not best & _XCD_BIT:
            best = min(near, key=timings.get)
        self.set_variant(best)
        self.tuning = timings
        self.tuned, self._tuned_key = True, key
        _tune_cache_store(key, best, timings)
        return step - self.step

    @_on_stream
    def advance(self, max_steps: int | None = None) -> list:
        """Run until the next admission point: at most `max_steps` steps (or
        the chunk cap), stopping early at stimulus steps and completions.
        Returns the tissues that finished."""
        finished = []
        if not self.sims:
            return finished
        if self.blocks_dirty:
            self._rebuild_blocks()
        self._apply_due_stims()
        # stimuli due now are applied; the launch stops before the next one
        target = self.step + (max_steps or self.chunk_cap)
        for slot, sim in self.sims.items():
            target = min(target, sim.step0 + sim.n_steps)
            ev = self.events[slot]
            if ev:
                target = min(target, sim.step0 + ev[0][0])
        n = max(1, target - self.step)
        if self.autotune and (not self.tuned or self._tune_key() != self._tuned_key):
            done = self._tune(n)
            self.step += done
            n -= done
        if n > 0:
            check(lib().fw_batch_run(self.handle, self.step, n, 1 if self.use_graph else 0,
                                     self._stream()), "fw_batch_run")
            self.step += n
        for slot, sim in list(self.sims.items()):
            if self.step - sim.step0 >= sim.n_steps:
                finished.append(self.retire(slot))
        self.done.extend(finished)
        return finished

    def run_all(self) -> list:
        while self.sims:
            self.advance()
        return self.done

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            self.stream.synchronize()  # no launch of this batch still reads its tables
            lib().fw_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class _ModelShim:
    def __init__(self, dt):
        self.dt = dt


class _TissueDesc(ctypes.Structure):
    _fields_ = [("u", ctypes.c_void_p), ("un", ctypes.c_void_p),
                ("state", ctypes.c_void_p * _native.MAX_STATE), ("code", ctypes.c_void_p),
                ("act", ctypes.c_void_p), ("probe", ctypes.c_void_p),
                ("ni", ctypes.c_int64), ("nj", ctypes.c_int64), ("n", ctypes.c_int64),
                ("d", ctypes.c_double * 4), ("act_thr", ctypes.c_double),
                ("stencil", ctypes.c_int32), ("step0", ctypes.c_int32),
                ("act_start", ctypes.c_int32), ("probe_idx", ctypes.c_int32),
                ("probe_len", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("state_alt", ctypes.c_void_p * _native.BATCH_ALT_STATE)]
