"""Trackers.

Mirrors ``core/tracker/tracker.py``, ``core/tracker/tracker_sequence.py`` and
the activation-time / point-probe trackers of ``cpuwave2D/tracker`` and
``cpuwave3D/tracker``.  A tracker runs when ``start_time <= t <= end_time``
and ``step % self.step == 0`` and sees the pre-step potential and the
post-ionic state (cardiac_model.py:173-178).

* ``ActivationTime2D/3DTracker`` runs inside the fused step kernel: it keeps
  the step index of the first crossing per node (int32 on the device) and
  converts it to the model time of that step on output.
* ``ActionPotential``, ``MultiVariable`` and ``Variable`` trackers are point
  probes written by the fused kernel into a device buffer.
* Any other ``Tracker`` subclass runs on the host with the state copied back
  for the steps its gate opens.
"""
from __future__ import annotations

import copy
from pathlib import Path

import numpy as np


class Tracker:
    """core/tracker/tracker.py:7-101"""

    def __init__(self):
        self.model = None
        self.file_name = "tracked_data"
        self.path = "."
        self.start_time = 0
        self.end_time = np.inf
        self.step = 1

    def initialize(self, model):
        raise NotImplementedError

    def _track(self):
        raise NotImplementedError

    def track(self):
        if self.start_time > self.model.t or self.model.t > self.end_time:
            return
        if self.model.step % self.step != 0:
            return
        self._track()

    def clone(self):
        return copy.deepcopy(self)

    def write(self):
        np.save(Path(self.path, self.file_name).with_suffix(".npy"), self.output)


class TrackerSequence:
    """core/tracker/tracker_sequence.py:1-64"""

    def __init__(self):
        self.sequence = []
        self.model = None

    def initialize(self, model):
        self.model = model
        for tracker in self.sequence:
            tracker.initialize(model)

    def add_tracker(self, tracker):
        self.sequence.append(tracker)

    def remove_trackers(self):
        self.sequence = []

    def tracker_next(self):
        for tracker in self.sequence:
            tracker.track()


class ActivationTime2DTracker(Tracker):
    """cpuwave2D/tracker/activation_time_2d_tracker.py:7-75:
    ``act_t = where((act_t < 0) & (u > threshold), t, act_t)`` over every
    node, ``act_t`` starting at -1."""

    def __init__(self):
        Tracker.__init__(self)
        self.threshold = -40
        self.file_name = "act_time_2d"
        self._steps = None
        self._t_of_step = None
        self._act_t_host = None

    def initialize(self, model):
        self.model = model
        self._steps = -np.ones(np.shape(model.u), dtype=np.int32)
        self._t_of_step = np.zeros(0)
        self._act_t_host = None

    def _steps_host(self) -> np.ndarray:
        if self._steps is None:
            self._steps = -np.ones(np.shape(self.model.u), dtype=np.int32)
        return self._steps

    def _store_steps(self, steps: np.ndarray, t_hist) -> None:
        self._steps = np.ascontiguousarray(steps, dtype=np.int32)
        self._t_of_step = np.asarray(t_hist, dtype=np.float64)
        self._act_t_host = None

    @property
    def act_t(self) -> np.ndarray:
        if self._act_t_host is None:
            steps = self._steps_host()
            out = -np.ones(steps.shape, dtype=np.float64)
            hit = steps >= 0
            if hit.any():
                out[hit] = self._t_of_step[steps[hit]]
            self._act_t_host = out
        return self._act_t_host

    @property
    def output(self):
        return self.act_t


class ActivationTime3DTracker(ActivationTime2DTracker):
    """cpuwave3D/tracker/activation_time_3d_tracker.py:7-15"""

    def __init__(self):
        super().__init__()
        self.file_name = "act_time_3d"


class _ProbeTracker(Tracker):
    """Point probes filled by the fused kernel."""

    def _probe_entries(self, model, state_names):
        raise NotImplementedError

    def _append(self, values: np.ndarray) -> None:
        raise NotImplementedError

    @staticmethod
    def _flat_cells(model, cell_ind) -> np.ndarray:
        cells = np.atleast_2d(np.asarray(cell_ind))
        shape = np.shape(model.cardiac_tissue.mesh)
        return np.ravel_multi_index(tuple(cells.T), shape)

    @staticmethod
    def _slot(var, state_names) -> int:
        if var == "u":
            return -1
        return state_names.index(var)


class ActionPotential2DTracker(_ProbeTracker):
    """cpuwave2D/tracker/action_potential_2d_tracker.py:7-65: records
    ``u[cell_ind]`` on every tracked step."""

    def __init__(self):
        Tracker.__init__(self)
        self.act_pot = []
        self.cell_ind = [1, 1]
        self.file_name = "act_pot"

    def initialize(self, model):
        self.model = model

    def _probe_entries(self, model, state_names):
        return [(int(c), -1) for c in self._flat_cells(model, self.cell_ind)]

    def _append(self, values):
        self.act_pot.append(np.array(values))

    def _track(self):
        cell_ind = tuple(np.atleast_2d(self.cell_ind).T)
        self.act_pot.append(self.model.u[cell_ind])

    @property
    def output(self):
        return np.squeeze(self.act_pot)


class ActionPotential3DTracker(ActionPotential2DTracker):
    pass


class MultiVariable2DTracker(_ProbeTracker):
    """cpuwave2D/tracker/multi_variable_2d_tracker.py:8-98"""

    def __init__(self):
        Tracker.__init__(self)
        self.var_list = []
        self.cell_ind = [1, 1]
        self.dir_name = "multi_vars"
        self.vars = {}

    def initialize(self, model):
        self.vars = {}
        self.model = model
        for var_ in self.var_list:
            if var_ not in self.model.__dict__:
                raise ValueError(f"Variable '{var_}' not found in model.")
            self.vars[var_] = []

    def _probe_entries(self, model, state_names):
        cells = self._flat_cells(model, self.cell_ind)
        self._ncell = len(cells)
        return [(int(c), self._slot(v, state_names)) for v in self.var_list for c in cells]

    def _append(self, values):
        n = self._ncell
        for k, var_ in enumerate(self.var_list):
            self.vars[var_].append(np.array(values[k * n:(k + 1) * n]))

    def _track(self):
        cell_ind = tuple(np.atleast_2d(self.cell_ind).T)
        for var_ in self.var_list:
            self.vars[var_].append(self.model.__dict__[var_][cell_ind])

    @property
    def output(self):
        return {var_: np.squeeze(self.vars[var_]) for var_ in self.var_list}

    def write(self):
        if not Path(self.path, self.dir_name).is_dir():
            Path(self.path, self.dir_name).mkdir(parents=True)
        for var_ in self.var_list:
            np.save(Path(self.path, self.dir_name, f"{var_}.npy"), self.output[var_])


class MultiVariable3DTracker(MultiVariable2DTracker):
    pass


class Variable2DTracker(MultiVariable2DTracker):
    """cpuwave2D/tracker/variable_2d_tracker.py:7-52"""

    def __init__(self):
        super().__init__()
        self.cell_ind = [1, 1]

    @property
    def var_name(self):
        return self.var_list[0]

    @var_name.setter
    def var_name(self, value):
        self.var_list = [value]

    @property
    def output(self):
        return self.vars[self.var_name]

    def write(self):
        if not Path(self.path, self.dir_name).exists():
            Path(self.path, self.dir_name).mkdir(parents=True)
        np.save(Path(self.path, self.dir_name, self.var_name).with_suffix(".npy"), self.output)


class Variable3DTracker(Variable2DTracker):
    pass


# ---------------------------------------------------------------- device trackers

class _DeviceTracker(Tracker):
    """A tracker evaluated by a kernel on the device, on the potential as it
    stands before each step its gate opens (after that step's stimuli, as
    the reference's trackers see ``model.u``), without copying the state to
    the host."""

    def _device_begin(self, model, eng) -> None:
        pass

    def _device_track(self, model, eng) -> None:
        raise NotImplementedError

    def _device_end(self, model, eng) -> None:
        pass


def _lat_layers(counts, steps, times, shape):
    """The reference's layer list (local_activation_time_2d_tracker.py:60-73)
    from every node's crossing calls: at each tracked call with crossings, a
    new layer opens when a crossing node already has a time in the last one,
    then the crossing nodes take that call's time in the last layer.
    counts[c]: crossings of node c; steps[c, k]: tracked-call index of its
    k-th crossing."""
    n = int(np.prod(shape))
    layers = [-np.ones(n)]
    if counts.sum() == 0:
        return [layers[0].reshape(shape)]
    nodes, calls = [], []
    for k in range(int(counts.max())):
        sel = np.flatnonzero(counts > k)
        nodes.append(sel)
        calls.append(steps[sel, k].astype(np.int64))
    nodes, calls = np.concatenate(nodes), np.concatenate(calls)
    order = np.argsort(calls, kind="stable")
    calls, nodes = calls[order], nodes[order]
    bounds = np.flatnonzero(np.diff(calls)) + 1
    for grp_calls, grp_nodes in zip(np.split(calls, bounds), np.split(nodes, bounds)):
        if np.any(layers[-1][grp_nodes] > -1):
            layers.append(-np.ones(n))
        layers[-1][grp_nodes] = times[int(grp_calls[0])]
    return [layer.reshape(shape) for layer in layers]


class LocalActivationTime2DTracker(_DeviceTracker):
    """cpuwave2D/tracker/local_activation_time_2d_tracker.py:7-103: every
    crossing of ``threshold`` by every node, in layers.  The device keeps
    each node's crossing steps (``capacity`` per node) and the layers are
    rebuilt on output by the reference's rule."""

    def __init__(self):
        Tracker.__init__(self)
        self.act_t = []
        self.threshold = -40
        self.file_name = "local_act_time_2d"
        self.capacity = 64
        self._times = []
        self._dev = None

    def initialize(self, model):
        self.model = model
        self.act_t = [-np.ones(np.shape(model.u))]
        self._times = []
        self._dev = None

    def _device_begin(self, model, eng):
        torch = eng.torch
        if self._dev is None:
            n = eng.N
            self._dev = dict(state=torch.zeros(n, dtype=torch.int32, device=eng.dev),
                             steps=torch.zeros(n * self.capacity, dtype=torch.int32,
                                               device=eng.dev),
                             overflow=torch.zeros(1, dtype=torch.int32, device=eng.dev))

    def _device_track(self, model, eng):
        from ._native import check, lib
        d = self._dev
        check(lib().fw_lat_update(eng.plan.handle, eng.u_dev().data_ptr(), d["state"].data_ptr(),
                                  d["steps"].data_ptr(), self.capacity, float(self.threshold),
                                  len(self._times), d["overflow"].data_ptr(),
                                  eng.torch.cuda.current_stream().cuda_stream), "fw_lat_update")
        self._times.append(model.t)

    def _device_end(self, model, eng):
        d = self._dev
        if int(d["overflow"].item()):
            raise RuntimeError(f"a node crossed the threshold more than capacity="
                               f"{self.capacity} times; raise tracker.capacity")
        state = d["state"].cpu().numpy()
        steps = d["steps"].cpu().numpy().reshape(-1, self.capacity)
        self.act_t = _lat_layers(state >> 1, steps, np.asarray(self._times, dtype=np.float64),
                                 np.shape(model.u))

    @property
    def output(self):
        return np.array(self.act_t)


class LocalActivationTime3DTracker(LocalActivationTime2DTracker):
    """cpuwave3D/tracker/local_activation_time_3d_tracker.py:7-15"""

    def __init__(self):
        super().__init__()
        self.file_name = "local_act_time_3d"


class Period2DTracker(_ProbeTracker):
    """cpuwave2D/tracker/period_2d_tracker.py:8-78: activation periods at the
    ``cell_ind`` detectors.  The fused kernel samples u at the detectors on
    every tracked step; the crossing rule of the local-activation-time
    tracker and the layer rule run on the host over those samples."""

    def __init__(self):
        Tracker.__init__(self)
        self.cell_ind = []
        self.threshold = -40
        self.file_name = "period"
        self._rows = []

    def initialize(self, model):
        self.model = model
        self._rows = []
        self._t0 = len(getattr(model, "_t_hist", []))

    def _probe_entries(self, model, state_names):
        return [(int(c), -1) for c in self._flat_cells(model, self.cell_ind)]

    def _append(self, values):
        self._rows.append(np.array(values, dtype=np.float64))

    def _track(self):
        cells = tuple(np.atleast_2d(self.cell_ind).T)
        self._rows.append(np.asarray(self.model.u[cells], dtype=np.float64))

    @property
    def act_t(self):
        from .model import _gate_open
        t_hist = self.model._t_hist
        times = [t for k, t in enumerate(t_hist) if k >= self._t0 and _gate_open(self, t, k)]
        n = len(np.atleast_2d(self.cell_ind))
        layers = [-np.ones(n)]
        activated = np.zeros(n, dtype=bool)
        for u, t in zip(self._rows, times):
            cross = (u >= self.threshold) & ~activated
            activated = np.where(cross, True, activated)
            activated = np.where((u < self.threshold) & activated, False, activated)
            if np.any(layers[-1][cross] > -1):
                layers.append(-np.ones(n))
            layers[-1] = np.where(cross, t, layers[-1])
        return layers

    @property
    def output(self):
        import pandas as pd
        lats = pd.DataFrame(np.array(self.act_t).T)
        return lats.apply(lambda row: np.diff(row[row != -1]), axis=1)

    def write(self):
        self.output.to_csv(Path(self.path, self.file_name).with_suffix(".csv"))


class Period3DTracker(Period2DTracker):
    """cpuwave3D/tracker/period_3d_tracker.py:4-10"""


class SpiralWaveCore2DTracker(_DeviceTracker):
    """cpuwave2D/tracker/spiral_wave_core_2d_tracker.py:10-105: spiral tips
    between the potential at the previous tracked step and the current one,
    found by a kernel over every interior cell; on 3D grids every slice
    along the third axis (spiral_wave_core_3d_tracker.py:18-31)."""

    _columns = ["x", "y", "time", "step"]

    def __init__(self):
        Tracker.__init__(self)
        self.threshold = 0.5
        self.file_name = "spiral_wave_core"
        self.sprial_wave_cores = []
        self.capacity = 1 << 16
        self._dev = None

    def initialize(self, model):
        self.model = model
        self.sprial_wave_cores = []
        self._dev = None

    def _device_begin(self, model, eng):
        torch = eng.torch
        if self._dev is None:
            self._dev = dict(u_prev=eng.u_dev().clone(),
                             tips=torch.zeros(5 * self.capacity, dtype=torch.float64,
                                              device=eng.dev),
                             count=torch.zeros(1, dtype=torch.int32, device=eng.dev))

    def _device_track(self, model, eng):
        import pandas as pd
        from ._native import check, lib
        d = self._dev
        u = eng.u_dev()
        check(lib().fw_spiral_tips(eng.plan.handle, d["u_prev"].data_ptr(), u.data_ptr(),
                                   float(self.threshold), d["tips"].data_ptr(), self.capacity,
                                   d["count"].data_ptr(),
                                   eng.torch.cuda.current_stream().cuda_stream),
              "fw_spiral_tips")
        d["u_prev"].copy_(u)
        count = int(d["count"].item())
        if count > self.capacity:
            raise RuntimeError(f"{count} spiral tips exceed capacity={self.capacity}")
        tips = d["tips"][:5 * count].cpu().numpy().reshape(count, 5)
        if count:
            tips = tips[np.lexsort((tips[:, 1], tips[:, 0], tips[:, 2]))]
        self._store(pd, tips, model)

    def _store(self, pd, tips, model):
        df = pd.DataFrame({"x": tips[:, 3], "y": tips[:, 4]})
        df["time"] = model.t
        df["step"] = model.step
        self.sprial_wave_cores.append(df)

    @property
    def output(self):
        import pandas as pd
        validated = [df for df in self.sprial_wave_cores if not df.empty]
        if not validated:
            return pd.DataFrame(columns=self._columns)
        return pd.concat(validated, ignore_index=True)

    def write(self):
        self.output.to_csv(Path(self.path, self.file_name).with_suffix(".csv"))


class SpiralWaveCore3DTracker(SpiralWaveCore2DTracker):
    """cpuwave3D/tracker/spiral_wave_core_3d_tracker.py:7-31: tips per slice
    k, with the slice in column z."""

    _columns = ["x", "y", "z", "time", "step"]

    def _store(self, pd, tips, model):
        # the reference appends one frame per slice, slices in order
        for k in range(np.shape(model.u)[2]):
            sel = tips[tips[:, 2] == k]
            df = pd.DataFrame({"x": sel[:, 3], "y": sel[:, 4]})
            df["z"] = k
            df["time"] = model.t
            df["step"] = model.step
            self.sprial_wave_cores.append(df)


class ECG2DTracker(_DeviceTracker):
    """cpuwave2D/tracker/ecg_2d_tracker.py:9-151: per measuring point, the
    sum over active nodes of (u_tr - u) / (d * dr), u_tr the diffused
    potential; one fused diffusion + reduction kernel per tracked step."""

    def __init__(self, measure_coords=None):
        Tracker.__init__(self)
        self.measure_coords = measure_coords
        self.ecg = []
        self.file_name = "ecg.npy"
        self._dev = None
        self._rows_cap = 4096

    def initialize(self, model):
        self.model = model
        self.measure_coords = np.atleast_2d(self.measure_coords)
        self.ecg = []
        self._dev = None

    def _device_begin(self, model, eng):
        torch = eng.torch
        if self._dev is None:
            coords = np.ascontiguousarray(np.atleast_2d(self.measure_coords), dtype=np.float64)
            n = len(coords)
            self._dev = dict(coords=torch.from_numpy(coords.reshape(-1)).to(eng.dev), n=n,
                             out=torch.zeros((self._rows_cap, n), dtype=torch.float64,
                                             device=eng.dev), rows=0)

    def _flush(self):
        d = self._dev
        if d["rows"]:
            self.ecg.extend(list(d["out"][:d["rows"]].cpu().numpy()))
            d["rows"] = 0

    def _device_track(self, model, eng):
        from ._native import check, lib
        d = self._dev
        if d["rows"] == self._rows_cap:
            self._flush()
        row = d["out"][d["rows"]]
        check(lib().fw_ecg(eng.plan.handle, eng.u_dev().data_ptr(), d["coords"].data_ptr(),
                           d["n"], float(model.dr), row.data_ptr(),
                           eng.torch.cuda.current_stream().cuda_stream), "fw_ecg")
        d["rows"] += 1

    def _device_end(self, model, eng):
        self._flush()

    @property
    def output(self):
        return np.array(self.ecg)

    def write(self):
        if not Path(self.path).exists():
            Path(self.path).mkdir(parents=True)
        np.save(Path(self.path).joinpath(self.file_name).with_suffix(".npy"), self.output)


class ECG3DTracker(ECG2DTracker):
    """cpuwave3D/tracker/ecg_3d_tracker.py:9-140"""
