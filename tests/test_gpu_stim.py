"""Every stimulus class of the reference, applied by the device schedule,
against the oracle.

model.run() resolves the stimulus rules (stim_sequence.py:65-77, stim.py:52-56)
into windows of steps and hands them to the step kernel, which applies the
stimuli of step s + 1 to the potential it writes at step s
(fw_plan_set_stims); the first step of each library call is stimulated by
fw_stim_apply_step.  The oracle applies the same stimuli on the host before
each step, as the reference does.  Aliev-Panfilov FP64 steps are exact, so the
potential and the activation map must match bit for bit.  The same runs with
the stimuli applied step by step from the host (the path taken when a
sequence does not fit the schedule) must give the same bits.
"""
from __future__ import annotations

import numpy as np
import pytest

import finitewave_amd as fw
from oracle.oracle import OracleSim, OStim, OTracker
from tests.scenarios import boxed_mesh


def _model(cls, tissue, stims, t_max, graphs=True, host=False, precision="fp64"):
    m = cls()
    m.dt, m.dr, m.t_max = 0.01, 0.25, t_max
    m.prog_bar = False
    m.precision = precision
    m.use_graphs = graphs
    m.cardiac_tissue = tissue
    seq = fw.StimSequence()
    for s in stims:
        seq.add_stim(s)
    m.stim_sequence = seq
    tr = fw.ActivationTime2DTracker() if tissue.mesh.ndim == 2 else fw.ActivationTime3DTracker()
    tr.threshold = 0.5
    ts = fw.TrackerSequence()
    ts.add_tracker(tr)
    m.tracker_sequence = ts
    if host:
        # every stimulus applied by its own launch before its step
        def no_schedule(eng, stims_, iters):
            eng.clear_stim_schedule()
            return False
        m._schedule_stims = no_schedule
    return m, tr


def _oracle(mesh, stims, t_max, special=None):
    sim = OracleSim("aliev_panfilov", mesh, 0.01, 0.25, special_boundaries=special)
    sim.stims.extend(stims)
    tr = OTracker("act_time", threshold=0.5)
    sim.trackers.append(tr)
    sim.run(t_max)
    return sim, tr


def _area_matrix(shape, coords):
    m = np.zeros(shape, dtype=np.uint8)
    m[tuple(np.asarray(coords).T)] = 1
    return m


def _stims_2d(mesh):
    shape = mesh.shape
    rng = np.random.RandomState(5)
    mat_c = (rng.random_sample(shape) < 0.3).astype(np.uint8)
    mat_v = np.zeros(shape)
    mat_v[15:20, 2:8] = 2.5  # > 0 selects
    area = fw.StimCurrentArea2D(0.2, 4.0, 0.5, u_max=None)
    area.add_stim_point(np.array([12, 12]), mesh, size=3.5)
    prod = [
        fw.StimVoltageCoord2D(0, 1.0, 1, 4, 1, shape[1] - 1),
        fw.StimCurrentCoord2D(0.5, 3.0, 0.6, 10, 16, 3, 9, u_max=0.8),
        # overlaps the box above: order within a step matters
        fw.StimCurrentMatrix2D(0.55, 2.0, 0.3, mat_c, u_max=0.9),
        fw.StimVoltageMatrix2D(2.0, 0.7, mat_v),
        area,
    ]
    orc = [
        OStim("voltage", 0, 1.0, (slice(1, 4), slice(1, shape[1] - 1))),
        OStim("current", 0.5, 3.0, (slice(10, 16), slice(3, 9)), duration=0.6, u_max=0.8),
        OStim("current", 0.55, 2.0, mat_c, duration=0.3, u_max=0.9),
        OStim("voltage", 2.0, 0.7, mat_v),
        OStim("current", 0.2, 4.0, _area_matrix(shape, area.coords), duration=0.5),
    ]
    return prod, orc


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [True, False])
def test_every_2d_stimulus_class_matches_oracle(gpu, graphs):
    mesh = boxed_mesh((26, 24), fibrosis=0.1)
    tissue = fw.CardiacTissue2D(list(mesh.shape))
    tissue.mesh = mesh.copy()
    prod, orc = _stims_2d(tissue.mesh)
    m, tr = _model(fw.AlievPanfilov2D, tissue, prod, 5.0, graphs=graphs)
    m.run()
    sim, otr = _oracle(tissue.mesh, orc, 5.0)
    assert m.step == sim.step
    assert np.array_equal(m.u, sim.u)
    assert np.array_equal(np.asarray(tr.output), otr.data)
    assert all(s.passed for s in prod)


@pytest.mark.gpu
def test_every_3d_stimulus_class_matches_oracle(gpu):
    mesh = boxed_mesh((12, 11, 10), fibrosis=0.1)
    shape = mesh.shape
    tissue = fw.CardiacTissue3D(list(shape))
    tissue.mesh = mesh.copy()
    rng = np.random.RandomState(6)
    mat_c = (rng.random_sample(shape) < 0.25).astype(np.uint8)
    mat_v = np.zeros(shape, dtype=np.uint8)
    mat_v[6:9, 2:5, 2:6] = 1
    mat_l = np.zeros(shape, dtype=np.uint8)
    mat_l[2:5, 6:9, 3:8] = 1
    area = fw.StimCurrentArea3D(0.3, 5.0, 0.4, u_max=0.95)
    area.add_stim_point(np.array([6, 5, 5]), tissue.mesh, size=2.5)
    values = list(np.linspace(0.1, 1.2, 40))
    prod = [
        fw.StimVoltageCoord3D(0, 1.0, 1, 3, 1, 10, 1, 9),
        fw.StimCurrentCoord3D(0.4, 2.5, 0.5, 4, 8, 2, 6, 2, 7, u_max=0.85),
        fw.StimCurrentMatrix3D(0.6, 3.0, 0.2, mat_c),
        fw.StimVoltageMatrix3D(1.5, 0.6, mat_v),
        area,
        fw.StimVoltageListMatrix3D(1.0, values, 0.3, mat_l),
    ]
    orc = [
        OStim("voltage", 0, 1.0, (slice(1, 3), slice(1, 10), slice(1, 9))),
        OStim("current", 0.4, 2.5, (slice(4, 8), slice(2, 6), slice(2, 7)), duration=0.5,
              u_max=0.85),
        OStim("current", 0.6, 3.0, mat_c, duration=0.2),
        OStim("voltage", 1.5, 0.6, mat_v),
        OStim("current", 0.3, 5.0, _area_matrix(shape, area.coords), duration=0.4, u_max=0.95),
        OStim("voltage", 1.0, 0.0, mat_l, duration=0.3, values=values),
    ]
    m, tr = _model(fw.AlievPanfilov3D, tissue, prod, 4.0)
    m.run()
    sim, otr = _oracle(tissue.mesh, orc, 4.0)
    assert np.array_equal(m.u, sim.u)
    assert np.array_equal(np.asarray(tr.output), otr.data)
    # the list stimulus consumed one value per application, as the oracle's
    assert prod[-1].step == orc[-1].count > 0


@pytest.mark.gpu
def test_stimulus_on_frozen_nodes_matches_oracle(gpu):
    """Special-boundary nodes conduct but are never updated; a stimulus still
    sets their potential (cardiac_tissue.py:57-61 with stim masks over
    mesh == 1)."""
    mesh = boxed_mesh((20, 18))
    special = np.zeros(mesh.shape, dtype=np.int8)
    special[5:9, 4:7] = 1
    tissue = fw.CardiacTissue2D(list(mesh.shape))
    tissue.mesh = mesh.copy()
    tissue.special_boundaries = special
    prod = [fw.StimCurrentCoord2D(0.1, 2.0, 0.3, 3, 10, 2, 9),
            fw.StimVoltageCoord2D(1.0, 0.9, 6, 8, 1, 17)]
    orc = [OStim("current", 0.1, 2.0, (slice(3, 10), slice(2, 9)), duration=0.3),
           OStim("voltage", 1.0, 0.9, (slice(6, 8), slice(1, 17)))]
    m, _ = _model(fw.AlievPanfilov2D, tissue, prod, 3.0)
    m.run()
    sim, _ = _oracle(tissue.mesh, orc, 3.0, special=special)
    assert np.array_equal(m.u, sim.u)
    assert np.array_equal(m.u_new, sim.u_new)


@pytest.mark.gpu
def test_stimulus_window_across_library_calls_matches_oracle(gpu):
    """A current stimulus whose window spans the step loop's chunk boundary
    (4096 steps per library call): the next call stimulates its first step
    itself, the kernels the rest."""
    mesh = boxed_mesh((16, 16))
    tissue = fw.CardiacTissue2D(list(mesh.shape))
    tissue.mesh = mesh.copy()
    prod = [fw.StimCurrentCoord2D(40.9, 0.5, 0.5, 1, 15, 1, 4)]
    orc = [OStim("current", 40.9, 0.5, (slice(1, 15), slice(1, 4)), duration=0.5)]
    m, tr = _model(fw.AlievPanfilov2D, tissue, prod, 42.0)
    m.run()
    sim, otr = _oracle(tissue.mesh, orc, 42.0)
    assert np.array_equal(m.u, sim.u)
    assert np.array_equal(np.asarray(tr.output), otr.data)


@pytest.mark.gpu
@pytest.mark.parametrize("cls,precision", [(fw.AlievPanfilov2D, "fp64"),
                                           (fw.AlievPanfilov2D, "fp32"),
                                           (fw.TP062D, "fp32")])
def test_scheduled_stimuli_equal_host_applied_stimuli(gpu, cls, precision):
    """The device schedule and one launch per stimulus per step give the same
    bits (FP32 included, where no oracle is bitwise)."""
    mesh = boxed_mesh((22, 20), fibrosis=0.05)
    amp = 1.0 if cls is fw.AlievPanfilov2D else 60.0
    runs = []
    for host in (False, True):
        tissue = fw.CardiacTissue2D(list(mesh.shape))
        tissue.mesh = mesh.copy()
        stims = [fw.StimCurrentCoord2D(0, amp * 2.0, 1.0, 1, 21, 1, 5),
                 fw.StimCurrentCoord2D(0.5, amp, 0.5, 8, 14, 8, 14, u_max=amp * 0.8)]
        m, tr = _model(cls, tissue, stims, 3.0, host=host, precision=precision)
        if cls is fw.TP062D:
            tr.threshold = -40.0
        m.run()
        runs.append((m.u.copy(), np.asarray(tr.output).copy()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
