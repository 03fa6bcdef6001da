"""Stimulation: voltage sets and current injections applied to the pre-step
potential.

Mirrors ``core/stimulation/*.py``, ``cpuwave2D/stimulation/*.py`` and
``cpuwave3D/stimulation/*.py``.  A stimulus with ``t >= stim.t`` that has not
passed is applied at the start of a step, then ``passed = t >= stim.t +
duration`` (stim_sequence.py:74-77, stim.py:52-56): voltage stimuli (duration
0) fire once; current stimuli add ``dt*I`` on every step up to and including
the first with ``t >= t0 + duration``.  Only nodes with mesh == 1 are touched.

On the GPU the run resolves those rules ahead of time into a device schedule
(``schedule_windows``): each stimulus applies on a window of steps, and the
kernel of step s applies the stimuli of step s + 1 as it writes the potential
step s + 1 reads, so stimulated stretches need no host round trip and replay
from captured graphs like any other.  ``_apply_device`` (one small kernel over
the stimulus box) remains for steps the host runs one at a time.
"""
from __future__ import annotations

import numpy as np

from . import _native


class Stim:
    """core/stimulation/stim.py:4-56"""

    def __init__(self, time, duration=0.0):
        self.t = time
        self.duration = duration
        self.passed = False

    def initialize(self, model):
        self.passed = False

    def update_status(self, model):
        self.passed = model.t >= self.t + self.duration

    def stimulate(self, model):
        raise NotImplementedError

    # device application: a list of fw_stim_desc
    def _descs(self, model, eng):
        raise NotImplementedError

    def _apply_device(self, model, eng) -> None:
        for d in self._descs(model, eng):
            eng.apply_stim(d)

    # device schedule
    def _sched_entry(self, model, eng, n_apps):
        """(fw_stim_desc, device value list or None) for ``n_apps``
        applications from the next one on; None if this stimulus has no
        single-descriptor form."""
        try:
            descs = self._descs(model, eng)
        except NotImplementedError:
            return None
        return (descs[0], None) if len(descs) == 1 else None

    def _host_advance(self, model) -> None:
        """The host bookkeeping of one application made on the device."""
        self.update_status(model)


class StimCurrent(Stim):
    def __init__(self, time, curr_value, duration):
        super().__init__(time, duration)
        self.curr_value = curr_value


class StimVoltage(Stim):
    def __init__(self, time, volt_value, duration=0.0):
        super().__init__(time, duration)
        self.volt_value = volt_value


class StimSequence:
    """core/stimulation/stim_sequence.py:1-77"""

    def __init__(self):
        self.sequence = []
        self.model = None

    def initialize(self, model):
        self.model = model
        for stim in self.sequence:
            stim.initialize(model)

    def add_stim(self, stim):
        self.sequence.append(stim)

    def remove_stim(self):
        self.sequence = []

    def stimulate_next(self):
        for stim in self.sequence:
            if self.model.t >= stim.t and not stim.passed:
                stim.stimulate(self.model)
                stim.update_status(self.model)


def schedule_windows(stims, t, dt, iters):
    """Steps (offsets from the current step) on which each stimulus applies
    over the next ``iters`` steps starting at time ``t``, with the time
    advanced as the step loop advances it (t += dt, cardiac_model.py:176):
    a stimulus applies while ``t >= stim.t`` and it has not passed, and
    passes once ``t >= stim.t + duration`` after an application
    (stim_sequence.py:74-77, stim.py:52-56).  Returns [(stim, lo, hi)] for
    the stimuli that apply, in sequence order; the applications of one
    stimulus are consecutive steps."""
    out = {}
    passed = [bool(s.passed) for s in stims]
    pending = sum(1 for p in passed if not p)
    q = 0
    while q < iters and pending:
        for k, st in enumerate(stims):
            if passed[k] or not t >= st.t:
                continue
            lo = out.get(k, (q, q))[0]
            out[k] = (lo, q)
            if t >= st.t + st.duration:
                passed[k] = True
                pending -= 1
        t = t + dt
        q += 1
    return [(stims[k], lo, hi) for k, (lo, hi) in sorted(out.items())]


def _box_desc(kind, box, shape, value, u_max=None, mask=None):
    d = _native.fw_stim_desc()
    d.kind = kind
    for a in range(3):
        if a < len(box):
            lo, hi, _ = box[a].indices(shape[a])
        else:
            lo, hi = 0, 1
        d.lo[a] = lo
        d.hi[a] = max(lo, hi)  # numpy's empty slices (hi < lo) select nothing
    d.mask = mask.data_ptr() if mask is not None else 0
    d.value = float(value)
    d.clamp = 1 if u_max is not None else 0
    d.u_max = float(u_max) if u_max is not None else 0.0
    return d


class _CoordMixin:
    def _box(self):
        raise NotImplementedError

    def _host_region(self, model):
        box = self._box()
        roi_mesh = model.cardiac_tissue.mesh[box]
        return box, roi_mesh == 1


class _MaskMixin:
    _mask_dev = None
    _mask_src = None

    def _mask(self, model, eng):
        m = np.ascontiguousarray(np.asarray(self.matrix) > 0, dtype=np.uint8)
        if self._mask_dev is None or self._mask_src is None or \
                not np.array_equal(self._mask_src, m):
            self._mask_src = m
            self._mask_dev = eng.torch.from_numpy(m).to(eng.dev)
        return self._mask_dev

    def _full_box(self, shape):
        return tuple(slice(0, n) for n in shape)


# ---------------------------------------------------------------- coordinates
class StimCurrentCoord2D(StimCurrent, _CoordMixin):
    """cpuwave2D/stimulation/stim_current_coord_2d.py:7-85"""

    def __init__(self, time, curr_value, duration, x1, x2, y1, y2, u_max=None):
        super().__init__(time, curr_value, duration)
        self.x1, self.x2, self.y1, self.y2 = x1, x2, y1, y2
        self.u_max = u_max

    def _box(self):
        return (slice(self.x1, self.x2), slice(self.y1, self.y2))

    def stimulate(self, model):
        box, mask = self._host_region(model)
        model.u[box][mask] += model.dt * self.curr_value
        if self.u_max is not None:
            u = model.u[box][mask]
            model.u[box][mask] = np.where(u > self.u_max, self.u_max, u)

    def _descs(self, model, eng):
        return [_box_desc(1, self._box(), eng.shape, model.dt * self.curr_value, self.u_max)]


class StimCurrentCoord3D(StimCurrentCoord2D):
    """cpuwave3D/stimulation/stim_current_coord_3d.py:7-90"""

    def __init__(self, time, curr_value, duration, x1, x2, y1, y2, z1, z2, u_max=None):
        StimCurrent.__init__(self, time, curr_value, duration)
        self.x1, self.x2, self.y1, self.y2, self.z1, self.z2 = x1, x2, y1, y2, z1, z2
        self.u_max = u_max

    def _box(self):
        return (slice(self.x1, self.x2), slice(self.y1, self.y2), slice(self.z1, self.z2))


class StimVoltageCoord2D(StimVoltage, _CoordMixin):
    """cpuwave2D/stimulation/stim_voltage_coord_2d.py:6-67"""

    def __init__(self, time, volt_value, x1, x2, y1, y2):
        super().__init__(time, volt_value)
        self.x1, self.x2, self.y1, self.y2 = x1, x2, y1, y2

    def _box(self):
        return (slice(self.x1, self.x2), slice(self.y1, self.y2))

    def stimulate(self, model):
        box, mask = self._host_region(model)
        model.u[box][mask] = self.volt_value

    def _descs(self, model, eng):
        return [_box_desc(0, self._box(), eng.shape, self.volt_value)]


class StimVoltageCoord3D(StimVoltageCoord2D):
    """cpuwave3D/stimulation/stim_voltage_coord_3d.py:6-68"""

    def __init__(self, time, volt_value, x1, x2, y1, y2, z1, z2):
        StimVoltage.__init__(self, time, volt_value)
        self.x1, self.x2, self.y1, self.y2, self.z1, self.z2 = x1, x2, y1, y2, z1, z2

    def _box(self):
        return (slice(self.x1, self.x2), slice(self.y1, self.y2), slice(self.z1, self.z2))


# ---------------------------------------------------------------- matrices
class StimCurrentMatrix2D(StimCurrent, _MaskMixin):
    """cpuwave2D/stimulation/stim_current_matrix_2d.py:7-63"""

    def __init__(self, time, curr_value, duration, matrix, u_max=None):
        super().__init__(time, curr_value, duration)
        self.matrix = matrix
        self.u_max = u_max

    def stimulate(self, model):
        mask = (self.matrix > 0) & (model.cardiac_tissue.mesh == 1)
        model.u[mask] += model.dt * self.curr_value
        if self.u_max is not None:
            model.u[mask] = np.where(model.u[mask] > self.u_max, self.u_max, model.u[mask])

    def _descs(self, model, eng):
        return [_box_desc(1, self._full_box(eng.shape), eng.shape, model.dt * self.curr_value,
                          self.u_max, self._mask(model, eng))]


class StimCurrentMatrix3D(StimCurrentMatrix2D):
    pass


class StimVoltageMatrix2D(StimVoltage, _MaskMixin):
    """cpuwave2D/stimulation/stim_voltage_matrix_2d.py:6-46"""

    def __init__(self, time, volt_value, matrix):
        super().__init__(time, volt_value)
        self.matrix = matrix

    def stimulate(self, model):
        mask = (self.matrix > 0) & (model.cardiac_tissue.mesh == 1)
        model.u[mask] = self.volt_value

    def _descs(self, model, eng):
        return [_box_desc(0, self._full_box(eng.shape), eng.shape, self.volt_value, None,
                          self._mask(model, eng))]


class StimVoltageMatrix3D(StimVoltageMatrix2D):
    pass


class StimVoltageListMatrix3D(StimVoltage, _MaskMixin):
    """cpuwave3D/stimulation/stim_voltage_list_matrix_3d.py:6-72: sets the
    potential to ``volt_values[k]`` on the k-th application."""

    def __init__(self, time, volt_values, duration, matrix):
        super().__init__(time, volt_values, duration)
        self.matrix = matrix
        self.step = 0

    def initialize(self, model):
        super().initialize(model)
        total_steps = int(self.duration / model.dt) + 1
        if len(self.volt_value) < total_steps:
            raise ValueError("The length of the voltage values array should be greater than "
                             "the total number of steps.")

    def stimulate(self, model):
        mask = (self.matrix > 0) & (model.cardiac_tissue.mesh == 1)
        model.u[mask] = self.volt_value[self.step]
        self.step += 1

    def _descs(self, model, eng):
        d = _box_desc(0, self._full_box(eng.shape), eng.shape, self.volt_value[self.step], None,
                      self._mask(model, eng))
        self.step += 1
        return [d]

    def _sched_entry(self, model, eng, n_apps):
        vals = np.asarray(self.volt_value, dtype=np.float64)[self.step:self.step + n_apps]
        if len(vals) < n_apps:
            # the reference's stimulate() indexes past the list on that step
            raise IndexError(f"StimVoltageListMatrix3D: {n_apps} applications from value "
                             f"{self.step} on, {len(np.asarray(self.volt_value))} values")
        d = _box_desc(0, self._full_box(eng.shape), eng.shape, 0.0, None,
                      self._mask(model, eng))
        return d, eng.torch.from_numpy(np.ascontiguousarray(vals)).to(eng.dev)

    def _host_advance(self, model) -> None:
        self.step += 1
        self.update_status(model)


# ---------------------------------------------------------------- areas
class StimCurrentArea2D(StimCurrent, _MaskMixin):
    """cpuwave2D/stimulation/stim_current_area_2d.py:7-94"""

    def __init__(self, time, curr_value, duration, coords=None, u_max=None):
        super().__init__(time, curr_value, duration)
        self.coords = coords
        self.u_max = u_max

    def add_stim_point(self, coord, mesh, size=None):
        self._coord = coord
        if size is None:
            self.coords = np.atleast_2d(coord)
            return
        tissue_points = np.argwhere(mesh == 1)
        dist = np.linalg.norm(tissue_points - coord, axis=1)
        self.coords = tissue_points[dist < size]

    def initialize(self, model):
        coords = np.atleast_2d(np.asarray(self.coords))
        mask = model.cardiac_tissue.mesh[tuple(coords.T)] == 1
        if mask.sum() == 0:
            raise ValueError("The specified area does not have healthy cells.")
        self._coords = coords[mask]
        self._mask_dev = None
        self._mask_src = None
        return super().initialize(model)

    def _mask(self, model, eng):
        # the area's nodes as a device mask, built once per initialize() (the
        # area is fixed there, stim_current_area_2d.py:67-74)
        if self._mask_dev is None or self._mask_src != tuple(eng.shape):
            m = np.zeros(tuple(eng.shape), dtype=np.uint8)
            m[tuple(self._coords.T)] = 1
            self._mask_src = tuple(eng.shape)
            self._mask_dev = eng.torch.from_numpy(m).to(eng.dev)
        return self._mask_dev

    def stimulate(self, model):
        inds = tuple(self._coords.T)
        model.u[inds] += model.dt * self.curr_value
        if self.u_max is not None:
            model.u[inds] = np.where(model.u[inds] > self.u_max, self.u_max, model.u[inds])

    def _descs(self, model, eng):
        return [_box_desc(1, self._full_box(eng.shape), eng.shape, model.dt * self.curr_value,
                          self.u_max, self._mask(model, eng))]


class StimCurrentArea3D(StimCurrentArea2D):
    pass
