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
