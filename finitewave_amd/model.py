"""CardiacModel: the step-loop driver with the reference's semantics.

Mirrors ``core/model/cardiac_model.py``.  Per step: stimuli -> diffusion ->
ionic -> trackers -> ``t += dt`` -> ``step += 1`` -> swap -> commands ->
saver -> termination.  Diffusion, ionic update and the activation-time /
point-probe trackers are one fused GPU kernel; stimuli are resolved into a
device schedule that the same kernel applies (each step's kernel stimulates
the potential it hands to the next step); stretches of steps with no
host-side event run back to back inside the library (optionally replayed
from a HIP graph); steps with a command, a state save or a host-only tracker
synchronise the host numpy view of the state first.
"""
from __future__ import annotations

import copy
import warnings

import numpy as np

from . import _native

_CHUNK = 4096  # steps per library call between host synchronisations


def _gate_open(tracker, t, step) -> bool:
    # tracker.py:77-81
    if tracker.start_time > t or t > tracker.end_time:
        return False
    return step % tracker.step == 0


def _saver_due(saver, model) -> bool:
    """Whether the state saver writes at the model's current time: the
    saver's own gate (StateSaver.due)."""
    return saver.due(model)


class CardiacModel:
    """Base class of the ionic models (cardiac_model.py:9-239)."""

    def __init__(self):
        self.meta = {}
        self.cardiac_tissue = None
        self.stim_sequence = None
        self.tracker_sequence = None
        self.command_sequence = None
        self.state_loader = None
        self.state_saver = None
        self.stencil = None
        self.diffusion_kernel = None
        self.ionic_kernel = None
        self.u = np.ndarray
        self.u_new = np.ndarray
        self.dt = 0.
        self.dr = 0.
        self.t_max = 0.
        self.t = 0
        self.step = 0
        self.D_model = 1.
        self.prog_bar = True
        self.npfloat = np.float64
        self.state_vars = []
        # extensions: "fp64" (reference arithmetic) or "fp32" (FP32 state,
        # TP06 calcium and ion concentrations in FP64); None follows npfloat
        self.precision = None
        self.use_graphs = True
        self._engine = None
        self._t_hist = []

    # -- reference API -------------------------------------------------------
    def initialize(self):
        # cardiac_model.py:91-117
        self.u = np.zeros_like(self.cardiac_tissue.mesh, dtype=self.npfloat)
        self.u_new = self.u.copy()
        self.step = 0
        self.t = 0
        self._t_hist = []
        self.compute_weights()
        self.diffusion_kernel = self.stencil.select_diffusion_kernel()
        if self.stim_sequence:
            self.stim_sequence.initialize(self)
        if self.tracker_sequence:
            self.tracker_sequence.initialize(self)
        if self.command_sequence:
            self.command_sequence.initialize(self)
        if self.state_loader:
            self.state_loader.initialize(self)
        if self.state_saver:
            self.state_saver.initialize(self)

    def compute_weights(self):
        """Active indexes, stencil choice and the device plan (the weights
        themselves are rebuilt per step from the mesh; ``self.weights``
        materialises the reference's array on demand)."""
        self.cardiac_tissue.compute_myo_indexes()
        if self.stencil is None:
            self.stencil = self.select_stencil(self.cardiac_tissue)
        self._weights_cache = None
        if self._engine is not None:
            self._engine.close()
            self._engine = None

    @property
    def weights(self):
        if getattr(self, "_weights_cache", None) is None:
            self._weights_cache = self.stencil.compute_weights(self, self.cardiac_tissue)
        return self._weights_cache

    @weights.setter
    def weights(self, value):
        self._weights_cache = value

    def run_diffusion_kernel(self):
        self.diffusion_kernel(self.u_new, self.u, self.weights,
                              self.cardiac_tissue.myo_indexes)

    def run_ionic_kernel(self):
        """One ionic sweep over numpy state through fw_ionic (the reference's
        per-kernel entry point; run() uses the fused kernel instead)."""
        from .engine import _stream
        torch = _native.require_gpu()
        dev = torch.device("cuda", torch.cuda.current_device())
        prec = self._precision_id()
        mid = self._model_id()
        rd = torch.float64 if prec == _native.PREC_FP64 else torch.float32
        un = torch.from_numpy(np.ascontiguousarray(self.u_new)).to(dev, rd)
        u = torch.from_numpy(np.ascontiguousarray(self.u)).to(dev, rd)
        names = self._state_call_order()
        st = []
        offs = [_native.state_offset(mid, prec, k) for k in range(len(names))]
        for k, n in enumerate(names):
            dt = torch.float64 if _native.state_is_f64(mid, prec, k) else torch.float32
            a = getattr(self, n)
            if offs[k]:  # stored as the difference from offs[k]
                a = np.asarray(a, dtype=np.float64) - offs[k]
            st.append(torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt))
        idx = torch.from_numpy(np.ascontiguousarray(self.cardiac_tissue.myo_indexes,
                                                    dtype=np.int64)).to(dev)
        import ctypes
        ptrs = (ctypes.c_void_p * len(st))(*[s.data_ptr() for s in st])
        params = self._param_values()
        par = (ctypes.c_double * len(params))(*params)
        _native.check(_native.lib().fw_ionic(mid, prec, un.data_ptr(), u.data_ptr(), ptrs,
                                             len(st), idx.data_ptr(), int(idx.numel()),
                                             float(self.dt), par, len(params), _stream()),
                      "fw_ionic")
        self.u_new[...] = un.cpu().numpy()
        for n, s, off in zip(names, st, offs):
            getattr(self, n)[...] = s.cpu().numpy().astype(np.float64) + off if off else s.cpu().numpy()

    def check_termination(self):
        # cardiac_model.py:191-203
        max_iters = int(np.ceil(self.t_max / self.dt))
        return (self.t > self.t_max) or (self.step > max_iters)

    def select_stencil(self, cardiac_tissue):
        raise NotImplementedError

    def clone(self):
        return copy.deepcopy(self)

    # -- hooks the model classes provide ----------------------------------------
    def _model_id(self) -> int:
        raise NotImplementedError

    def _param_values(self) -> list:
        raise NotImplementedError

    def _state_call_order(self) -> list:
        raise NotImplementedError

    def _precision_id(self) -> int:
        p = self.precision
        if p is None:
            p = "fp32" if np.dtype(self.npfloat) == np.float32 else "fp64"
        if p not in ("fp64", "fp32"):
            raise ValueError(f"precision must be 'fp64' or 'fp32', not {p!r}")
        return _native.PREC_FP64 if p == "fp64" else _native.PREC_FP32

    # -- run ------------------------------------------------------------------
    def run(self, initialize=True, num_of_theads=None):
        """cardiac_model.py:130-189 on the GPU.  ``num_of_theads`` is accepted
        for compatibility and ignored (there is no CPU thread pool)."""
        from .engine import Engine
        if initialize:
            self.initialize()
        if self.t_max < self.t:
            raise ValueError("t_max must be greater than current t.")
        if self.state_loader:
            self.state_loader.load()
        iters = int(np.ceil((self.t_max - self.t) / self.dt))
        if self._engine is None:
            self._engine = Engine(self)
        eng = self._engine
        eng.upload()
        try:
            self._loop(eng, iters)
        finally:
            eng.download()
            self._finish_trackers(eng)

    # -- the step loop -----------------------------------------------------------
    def _loop(self, eng, iters: int) -> None:
        from .engine import Probes
        from .tracker import ActivationTime2DTracker, _DeviceTracker, _ProbeTracker
        stims = list(self.stim_sequence.sequence) if self.stim_sequence else []
        trackers = list(self.tracker_sequence.sequence) if self.tracker_sequence else []
        commands = list(self.command_sequence.sequence) if self.command_sequence else []
        act_trackers = [t for t in trackers if isinstance(t, ActivationTime2DTracker)]
        probe_trackers = [t for t in trackers if isinstance(t, _ProbeTracker)]
        dev_trackers = [t for t in trackers if isinstance(t, _DeviceTracker)]
        host_trackers = [t for t in trackers
                         if not isinstance(t, (ActivationTime2DTracker, _ProbeTracker,
                                               _DeviceTracker))]
        for tr in dev_trackers:
            tr._device_begin(self, eng)
        torch = eng.torch
        # activation-time maps live on the device during the run
        for tr in act_trackers:
            eng.act[tr] = torch.from_numpy(tr._steps_host()).to(eng.dev)
        fused_act = act_trackers[0] if act_trackers else None
        extra_act = act_trackers[1:]
        # point probes: one device list for all probe trackers
        probes = None
        entries, slices = [], []
        for tr in probe_trackers:
            e = tr._probe_entries(self, eng.state_names)
            slices.append((len(entries), len(e)))
            entries.extend(e)
        if entries:
            eng.set_probe_bits(sorted({c for c, _ in entries}))
            probes = Probes(eng, entries, _CHUNK)
        else:
            eng.set_probe_bits([])
        bar = None
        if self.prog_bar:
            try:
                from tqdm import tqdm
                bar = tqdm(total=iters, desc=f"Running {self.__class__.__name__}")
            except Exception:  # pragma: no cover
                bar = None

        chunk_flags, chunk_rows, chunk_row_owners = [], [], []
        chunk_start = [self.step]

        def flush():
            n = len(chunk_flags)
            if n == 0:
                return
            s0 = chunk_start[0]
            flags = np.array(chunk_flags, dtype=np.uint8) if fused_act is not None else None
            rows = np.array(chunk_rows, dtype=np.int32) if probes is not None else None
            eng.run_steps(s0, n, fused_act, flags, probes, rows, use_graph=self.use_graphs)
            if probes is not None and any(r >= 0 for r in chunk_rows):
                self._collect_probes(probes, probe_trackers, slices, chunk_row_owners)
            if bar is not None:
                bar.update(n)
            chunk_flags.clear()
            chunk_rows.clear()
            chunk_row_owners.clear()

        dt = self.dt
        # stimuli on the device schedule: their steps need no host event
        sched = self._schedule_stims(eng, stims, iters)
        for _ in range(iters):
            t, step = self.t, self.step
            if len(chunk_flags) == 0:
                chunk_start[0] = step
            due_stims = [s for s in stims if t >= s.t and not s.passed]
            act_open = fused_act is not None and _gate_open(fused_act, t, step)
            extra_open = [tr for tr in extra_act if _gate_open(tr, t, step)]
            probe_open = [tr for tr in probe_trackers if _gate_open(tr, t, step)]
            host_open = [tr for tr in host_trackers if _gate_open(tr, t, step)]
            dev_open = [tr for tr in dev_trackers if _gate_open(tr, t, step)]
            t_next = t + dt
            cmd_due = any((not c.passed) and t_next >= c.t for c in commands)
            saver_due = False
            if self.state_saver:
                t_keep = self.t
                self.t = t_next
                saver_due = _saver_due(self.state_saver, self)
                self.t = t_keep
            plain = not ((due_stims and not sched) or extra_open or host_open or dev_open or
                         cmd_due or saver_due)
            # a plain step's probe row is its place in the chunk (a host-event
            # step runs alone, in row 0)
            row = len(chunk_flags) if probe_open else -1
            self._t_hist.append(t)
            if plain:
                for s in due_stims:
                    s._host_advance(self)
                chunk_flags.append(1 if act_open else 0)
                chunk_rows.append(row)
                chunk_row_owners.append(probe_open)
                self.t = t_next
                self.step = step + 1
                if len(chunk_flags) >= _CHUNK:
                    flush()
                if self.check_termination():
                    break
                continue
            # a step with host-side work
            flush()
            chunk_start[0] = step
            if sched:
                eng.apply_stims_at(step)
                for s in due_stims:
                    s._host_advance(self)
            else:
                for s in due_stims:
                    s._apply_device(self, eng)
                    s.update_status(self)
            # device trackers see the pre-step potential after this step's stimuli
            for tr in dev_open:
                tr._device_track(self, eng)
            chunk_flags.append(1 if act_open else 0)
            chunk_rows.append(0 if probe_open else -1)
            chunk_row_owners.append(probe_open)
            if host_open:
                # host trackers see the pre-step u and the post-ionic state
                self._run_single(eng, chunk_flags, chunk_rows, chunk_row_owners, fused_act,
                                 probes, probe_trackers, slices, pre_swap_hook=True)
                for tr in extra_open:
                    eng.act_update(tr, step)
                self._sync_pre_swap(eng)
                for tr in host_open:
                    tr._track()
            else:
                self._run_single(eng, chunk_flags, chunk_rows, chunk_row_owners, fused_act,
                                 probes, probe_trackers, slices)
                for tr in extra_open:
                    eng.act_update(tr, step)
            if bar is not None:
                bar.update(1)
            self.t = t_next
            self.step = step + 1
            eng.swapped = not eng.swapped
            if cmd_due:
                eng.download()
                for c in commands:
                    if not c.passed and c.update_status(self):
                        c.execute(self)
                self._after_host_change(eng)
            if self.state_saver and _saver_due(self.state_saver, self):
                eng.download()
                self.state_saver.save()
            if self.check_termination():
                if self.state_saver and _saver_due(self.state_saver, self):
                    eng.download()
                    self.state_saver.save()
                break
        flush()
        if sched:
            eng.clear_stim_schedule()
        if bar is not None:
            bar.close()

    def _schedule_stims(self, eng, stims, iters) -> bool:
        """Puts the run's stimuli on the device schedule (fw_plan_set_stims):
        the steps each applies on, from the time fold, and its descriptor.
        False (host-applied stimuli) when there are more than the schedule
        holds or one has no descriptor form."""
        from .stimulation import schedule_windows
        eng.clear_stim_schedule()
        if not stims or len(stims) > _native.MAX_STIMS:
            return False
        entries = []
        for st, lo, hi in schedule_windows(stims, self.t, self.dt, iters):
            e = st._sched_entry(self, eng, hi - lo + 1)
            if e is None:
                return False
            entries.append((e[0], e[1], self.step + lo, self.step + hi))
        if not entries:
            return False
        eng.set_stim_schedule(entries)
        return True

    def _run_single(self, eng, flags, rows, owners, fused_act, probes, probe_trackers, slices,
                    pre_swap_hook=False):
        """One fused step from the current orientation without the swap
        bookkeeping (the caller swaps)."""
        s0 = self.step
        f = np.array(flags, dtype=np.uint8) if fused_act is not None else None
        r = np.array(rows, dtype=np.int32) if probes is not None else None
        eng.run_steps(s0, 1, fused_act, f, probes, r, use_graph=False, stim_first=False)
        eng.swapped = not eng.swapped  # run_steps swapped; undo, caller owns it
        if probes is not None and rows[0] >= 0:
            self._collect_probes(probes, probe_trackers, slices, owners)
        flags.clear()
        rows.clear()
        owners.clear()

    def _sync_pre_swap(self, eng) -> None:
        """Host view during a step: u = pre-step potential, state after the
        ionic update (what the reference's trackers see)."""
        m = self
        m.u = eng.u_dev().cpu().numpy()
        m.u_new = eng.u_new_dev().cpu().numpy()
        for name, a in eng.state_values():
            setattr(m, name, a)

    def _after_host_change(self, eng) -> None:
        """A command may rewrite state, mesh, fibers or weights.  A command
        that calls ``compute_weights()`` (the reference's way to make a mesh
        change take effect) drops the model's engine; the running loop's
        engine is rebuilt from the new mesh and kept as the model's."""
        eng.build()
        eng.upload()
        self._engine = eng

    def _collect_probes(self, probes, probe_trackers, slices, owners) -> None:
        data = probes.buf[: len(owners)].cpu().numpy()
        for r, open_trackers in enumerate(owners):
            for tr in open_trackers:
                k = probe_trackers.index(tr)
                a, n = slices[k]
                tr._append(data[r, a:a + n])

    def _finish_trackers(self, eng) -> None:
        from .tracker import _DeviceTracker
        for tr, dev_steps in eng.act.items():
            tr._store_steps(dev_steps.cpu().numpy(), self._t_hist)
        eng.act = {}
        trackers = list(self.tracker_sequence.sequence) if self.tracker_sequence else []
        for tr in trackers:
            if isinstance(tr, _DeviceTracker) and getattr(tr, "_dev", None) is not None:
                tr._device_end(self, eng)
