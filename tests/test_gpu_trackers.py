"""Device trackers: local activation time, period, spiral-wave core, ECG.

Each is checked twice: against the reference's own known-answer tests
(tests/test_trackers_2d.py, tests/test_trackers_3d.py of the reference, the
same fixtures) and against the oracle's restatement of the tracker on the
same run (FP64 path: the potentials are the oracle's to 1e-12, bitwise for
Aliev-Panfilov and Barkley, so crossings, layers and tips coincide; ECG sums
differ only by summation order)."""
from __future__ import annotations

import numpy as np
import pytest

import finitewave_amd as fw
from oracle.oracle import OracleSim, OStim, OTracker
from tests.scenarios import boxed_mesh


def _spiral_model(dim=2):
    # spiral_model fixture (test_trackers_2d.py:25-39 / test_trackers_3d.py:28-44)
    shape = [100, 100] if dim == 2 else [100, 100, 3]
    tissue = fw.CardiacTissue2D(shape) if dim == 2 else fw.CardiacTissue3D(shape)
    seq = fw.StimSequence()
    if dim == 2:
        seq.add_stim(fw.StimVoltageCoord2D(0, 1, 0, 100, 0, 3))
        seq.add_stim(fw.StimVoltageCoord2D(5, 1, 0, 50, 0, 100))
        model = fw.Barkley2D()
    else:
        seq.add_stim(fw.StimVoltageCoord3D(0, 1, 0, 100, 0, 3, 0, 3))
        seq.add_stim(fw.StimVoltageCoord3D(5, 1, 0, 50, 0, 100, 0, 3))
        model = fw.Barkley3D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 20
    model.cardiac_tissue = tissue
    model.stim_sequence = seq
    model.prog_bar = False
    return model


def _oracle_spiral(dim=2):
    shape = (100, 100) if dim == 2 else (100, 100, 3)
    sim = OracleSim("barkley", boxed_mesh(shape), 0.01, 0.25)
    ext = (slice(0, 3),) if dim == 3 else ()
    sim.stims.append(OStim("voltage", 0, 1, (slice(0, 100), slice(0, 3)) + ext))
    sim.stims.append(OStim("voltage", 5, 1, (slice(0, 50), slice(0, 100)) + ext))
    return sim


def _run(model, tracker):
    ts = fw.TrackerSequence()
    ts.add_tracker(tracker)
    model.tracker_sequence = ts
    model.run()
    return tracker


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_period_tracker(gpu, dim):
    # test_spiral_wave_period_2d_tracker / _3d_tracker: mean 3.5 +- 0.2
    cells = [[80, 80], [20, 70], [40, 10], [25, 90]]
    if dim == 3:
        cells = [c + [1] for c in cells]
    tr = (fw.Period2DTracker if dim == 2 else fw.Period3DTracker)()
    tr.cell_ind = np.array(cells)
    tr.threshold, tr.start_time, tr.step = 0.5, 10, 10
    _run(_spiral_model(dim), tr)
    periods = tr.output
    mean = np.mean(np.array([np.mean(x) if len(x) > 0 else np.nan for x in periods]))
    assert abs(mean - 3.5) <= 0.2
    sim = _oracle_spiral(dim)
    otr = OTracker("period", threshold=0.5, cells=cells, start_time=10, step=10)
    sim.trackers.append(otr)
    sim.run(20)
    ref = otr.data["act_t"]
    assert len(tr.act_t) == len(ref)
    for a, b in zip(tr.act_t, ref):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_spiral_wave_core_tracker(gpu, dim):
    # test_spiral_wave_core_2d_tracker: tips within x 32..38, y 47..53
    tr = (fw.SpiralWaveCore2DTracker if dim == 2 else fw.SpiralWaveCore3DTracker)()
    tr.threshold, tr.start_time, tr.step = 0.5, 12, 10
    _run(_spiral_model(dim), tr)
    out = tr.output
    x, y = out["x"].to_numpy(), out["y"].to_numpy()
    assert len(x) > 0
    assert x.min() >= 32 and x.max() <= 38 and y.min() >= 47 and y.max() <= 53
    sim = _oracle_spiral(dim)
    otr = OTracker("spiral", threshold=0.5, start_time=12, step=10)
    sim.trackers.append(otr)
    sim.run(20)
    ref = np.array(otr.data["tips"])
    cols = ["x", "y", "time", "step"] if dim == 2 else ["x", "y", "z", "time", "step"]
    got = out[cols].to_numpy(dtype=np.float64)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_ecg_tracker(gpu, dim):
    # planar_model + test_ecg_2d_tracker / test_ecg_3d_tracker
    shape = [50, 5] if dim == 2 else [50, 5, 3]
    tissue = fw.CardiacTissue2D(shape) if dim == 2 else fw.CardiacTissue3D(shape)
    seq = fw.StimSequence()
    if dim == 2:
        seq.add_stim(fw.StimCurrentCoord2D(5, 5, 0.5, 0, 5, 0, 5))
    else:
        seq.add_stim(fw.StimCurrentCoord3D(5, 5, 0.5, 0, 5, 0, 5, 0, 3))
    model = (fw.AlievPanfilov2D if dim == 2 else fw.AlievPanfilov3D)()
    model.dt, model.dr, model.t_max = 0.0015, 0.25, 15
    model.cardiac_tissue = tissue
    model.stim_sequence = seq
    model.prog_bar = False
    coords = np.array([[25, 2, 0 if dim == 2 else 1], [10, 1, 3], [40, 4, 2]])
    tr = (fw.ECG2DTracker if dim == 2 else fw.ECG3DTracker)()
    tr.start_time, tr.step, tr.measure_coords = 0, 10, coords
    _run(model, tr)
    ecg = tr.output
    e0 = ecg.T[0]
    assert e0.max() > 0.001 and e0.min() < -0.001 and np.argmax(e0) > 100
    sim = OracleSim("aliev_panfilov", boxed_mesh(tuple(shape)), 0.0015, 0.25)
    box = (slice(0, 5), slice(0, 5)) + ((slice(0, 3),) if dim == 3 else ())
    sim.stims.append(OStim("current", 5, 5, box, duration=0.5))
    otr = OTracker("ecg", step=10, coords=coords)
    sim.trackers.append(otr)
    sim.run(15)
    ref = np.array(otr.data)
    assert ecg.shape == ref.shape
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(ecg - ref)) <= 1e-12 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_local_activation_time_tracker(gpu, precision):
    # test_local_activation_time_2d_tracker: cable, second stimulus at 45
    tissue = fw.CardiacTissue2D([12, 3])
    seq = fw.StimSequence()
    seq.add_stim(fw.StimCurrentCoord2D(0, 5, 0.5, 0, 5, 0, 3))
    seq.add_stim(fw.StimVoltageCoord2D(45, 1, 0, 5, 0, 10))
    model = fw.AlievPanfilov2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 50
    model.precision = precision
    model.cardiac_tissue = tissue
    model.stim_sequence = seq
    model.prog_bar = False
    tr = fw.LocalActivationTime2DTracker()
    tr.threshold, tr.step, tr.start_time = 0.5, 1, 0
    _run(model, tr)
    lats = tr.output
    assert len(lats) == 2
    lat1, lat2 = lats[:, 10, 1]
    assert lat1 < lat2
    assert 1.5 <= 5 * 0.25 / lat1 <= 2
    assert 1.5 <= 5 * 0.25 / (lat2 - 45) <= 2
    if precision == "fp64":
        sim = OracleSim("aliev_panfilov", boxed_mesh((12, 3)), 0.01, 0.25)
        sim.stims.append(OStim("current", 0, 5, (slice(0, 5), slice(0, 3)), duration=0.5))
        sim.stims.append(OStim("voltage", 45, 1, (slice(0, 5), slice(0, 10))))
        otr = OTracker("lat", threshold=0.5)
        sim.trackers.append(otr)
        sim.run(50)
        ref = np.array(otr.data["act_t"])
        assert np.array_equal(lats, ref)


@pytest.mark.gpu
def test_local_activation_time_3d_spiral_matches_oracle(gpu):
    """Many re-entries: the 3D Barkley spiral tracked every 5 steps."""
    model = _spiral_model(3)
    tr = fw.LocalActivationTime3DTracker()
    tr.threshold, tr.step = 0.5, 5
    _run(model, tr)
    sim = _oracle_spiral(3)
    otr = OTracker("lat", threshold=0.5, step=5)
    sim.trackers.append(otr)
    sim.run(20)
    ref = np.array(otr.data["act_t"])
    got = tr.output
    assert got.shape == ref.shape and len(ref) >= 3
    assert np.array_equal(got, ref)
