"""The finitewave API on the GPU path, checked against the oracle and the
reference's own known-answer windows (tests/test_models_2d.py,
tests/test_trackers_2d.py, tests/test_basics.py of the reference)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle.oracle import OracleSim, OStim, OTracker
from tests.scenarios import boxed_mesh, make_pair, random_fibers
from tests.test_oracle import WINDOWS, _apd

import finitewave_amd as fw

CLASSES2D = {"aliev_panfilov": fw.AlievPanfilov2D, "barkley": fw.Barkley2D,
             "mitchell_schaeffer": fw.MitchellSchaeffer2D, "fenton_karma": fw.FentonKarma2D,
             "bueno_orovio": fw.BuenoOrovio2D, "luo_rudy91": fw.LuoRudy912D,
             "tp06": fw.TP062D, "courtemanche": fw.Courtemanche2D}


def _protocol_model(cls, cur, dur, t_calc, t_pre, dim=2):
    # tests/test_models_2d.py:8-72
    shape = [5, 3] if dim == 2 else [5, 3, 3]
    tissue = fw.CardiacTissue2D(shape) if dim == 2 else fw.CardiacTissue3D(shape)
    seq = fw.StimSequence()
    for k in range(4):
        if dim == 2:
            seq.add_stim(fw.StimCurrentCoord2D(k * t_pre, cur, dur, 0, 2, 0, 3))
        else:
            seq.add_stim(fw.StimCurrentCoord3D(k * t_pre, cur, dur, 0, 2, 0, 3, 0, 3))
    model = cls()
    model.dt, model.dr = 0.01, 0.25
    model.t_max = 3 * t_pre + t_calc
    model.cardiac_tissue = tissue
    model.stim_sequence = seq
    model.prog_bar = False
    tr = fw.ActionPotential2DTracker() if dim == 2 else fw.ActionPotential3DTracker()
    tr.cell_ind = [3, 1] if dim == 2 else [3, 1, 1]
    tr.step = 1
    ts = fw.TrackerSequence()
    ts.add_tracker(tr)
    model.tracker_sequence = ts
    return model, tr


@pytest.mark.gpu
@pytest.mark.parametrize("case", WINDOWS, ids=[w[0] for w in WINDOWS])
def test_reference_windows_on_gpu(gpu, case):
    name, cur, dur, t_calc, t_pre, thr, (lo, hi), amp = case
    model, tr = _protocol_model(CLASSES2D[name], cur, dur, t_calc, t_pre)
    model.run()
    u = tr.output
    apd = _apd(u, 0.01, thr)
    assert apd is not None and lo <= apd <= hi, apd
    if amp[0] == "approx_max":
        assert abs(np.max(u) - amp[1]) <= amp[2] and abs(np.min(u)) <= 0.01
    else:
        assert np.max(u) > amp[1] and np.min(u) < amp[2]
    # and the whole trace follows the oracle's
    from tests.test_oracle import run_reference_protocol
    ref = run_reference_protocol(name, cur, dur, t_calc, t_pre)
    assert u.shape == ref.shape
    scale = max(1.0, float(np.max(np.abs(ref))))
    assert float(np.max(np.abs(u - ref))) <= 1e-9 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("cls", [fw.AlievPanfilov3D, fw.TP063D])
def test_reference_windows_3d_on_gpu(gpu, cls):
    name = "aliev_panfilov" if cls is fw.AlievPanfilov3D else "tp06"
    _, cur, dur, t_calc, t_pre, thr, (lo, hi), _ = [w for w in WINDOWS if w[0] == name][0]
    model, tr = _protocol_model(cls, cur, dur, t_calc, t_pre, dim=3)
    model.run()
    apd = _apd(tr.output, 0.01, thr)
    assert lo <= apd <= hi


def _cable(t_max=3.0):
    # tests/test_trackers_2d.py:7-22
    tissue = fw.CardiacTissue2D([12, 3])
    seq = fw.StimSequence()
    seq.add_stim(fw.StimCurrentCoord2D(0, 5, 0.5, 0, 5, 0, 3))
    model = fw.AlievPanfilov2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, t_max
    model.cardiac_tissue = tissue
    model.stim_sequence = seq
    model.prog_bar = False
    return model


@pytest.mark.gpu
def test_activation_time_cable_velocity(gpu):
    # tests/test_trackers_2d.py:123-146
    model = _cable()
    tr = fw.ActivationTime2DTracker()
    tr.threshold = 0.5
    ts = fw.TrackerSequence()
    ts.add_tracker(tr)
    model.tracker_sequence = ts
    model.run()
    speed = 5 * model.dr / tr.output[10, 1]
    assert 1.5 <= speed <= 2
    sim = OracleSim("aliev_panfilov", model.cardiac_tissue.mesh, 0.01, 0.25)
    sim.stims.append(OStim("current", 0, 5, (slice(0, 5), slice(0, 3)), duration=0.5))
    ot = OTracker("act_time", threshold=0.5)
    sim.trackers.append(ot)
    sim.run(3.0)
    assert np.array_equal(tr.output, ot.data)


@pytest.mark.gpu
def test_multivariable_tracker_v(gpu):
    # tests/test_trackers_2d.py:183-202
    model = _cable(30)
    tr = fw.MultiVariable2DTracker()
    tr.cell_ind = [10, 1]
    tr.var_list = ["v"]
    ts = fw.TrackerSequence()
    ts.add_tracker(tr)
    model.tracker_sequence = ts
    model.run()
    v = tr.output["v"]
    assert abs(np.max(v) - 2) <= 0.1


@pytest.mark.gpu
def test_planar_wave_activation_map_and_state_match_oracle(gpu):
    # the 102x102 Aliev-Panfilov configuration, 1,000 steps, activation map
    n = 102
    tissue = fw.CardiacTissue2D([n, n])
    model = fw.AlievPanfilov2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 10
    model.cardiac_tissue = tissue
    model.prog_bar = False
    seq = fw.StimSequence()
    seq.add_stim(fw.StimVoltageCoord2D(0, 1.0, 0, n, 0, 6))
    model.stim_sequence = seq
    tr = fw.ActivationTime2DTracker()
    tr.threshold = 0.5
    ts = fw.TrackerSequence()
    ts.add_tracker(tr)
    model.tracker_sequence = ts
    model.run()
    assert model.step == 1000
    sim = OracleSim("aliev_panfilov", tissue.mesh, 0.01, 0.25)
    sim.stims.append(OStim("voltage", 0, 1.0, (slice(0, n), slice(0, 6))))
    ot = OTracker("act_time", threshold=0.5)
    sim.trackers.append(ot)
    sim.run(10)
    assert np.array_equal(model.u, sim.u)
    assert np.array_equal(model.v, sim.state["v"])
    assert np.array_equal(tr.output, ot.data)


@pytest.mark.gpu
def test_graph_replay_equals_plain_launches(gpu):
    outs = []
    for graphs in (True, False):
        shape = (40, 44)
        model, _ = make_pair("tp06", 2, shape, stencil="aniso", fibers=random_fibers(shape))
        model.use_graphs = graphs
        model.t_max = 5.0
        tr = fw.ActivationTime2DTracker()
        tr.threshold = -40
        ts = fw.TrackerSequence()
        ts.add_tracker(tr)
        model.tracker_sequence = ts
        tr.initialize(model)
        model.run(initialize=False)
        outs.append((model.u.copy(), model.casr.copy(), tr.output.copy()))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_commands(gpu):
    # tests/test_basics.py:52-76
    n = 5
    model = fw.FentonKarma2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 10
    model.prog_bar = False

    class ExcitationCommand(fw.Command):
        def execute(self, model):
            model.u[1:-1, 1:-1] = 1

    cs = fw.CommandSequence()
    cs.add_command(ExcitationCommand(5))
    model.cardiac_tissue = fw.CardiacTissue2D([n, n])
    model.stim_sequence = fw.StimSequence()
    model.command_sequence = cs
    model.run()
    assert np.mean(model.u[1:-1, 1:-1]) > 0.5


@pytest.mark.gpu
def test_state_save_and_load(gpu, tmp_path):
    # tests/test_basics.py:8-50
    n = 5
    tissue = fw.CardiacTissue2D([n, n])
    seq = fw.StimSequence()
    seq.add_stim(fw.StimCurrentCoord2D(0, 10, 0.5, 1, n // 2, n // 2 + 1, n // 2, n // 2 + 1))
    saver = fw.StateSaverCollection()
    saver.savers.append(fw.StateSaver(str(tmp_path / "state_0"), time=3))
    model = fw.FentonKarma2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 5
    model.cardiac_tissue = tissue
    model.stim_sequence = seq
    model.state_saver = saver
    model.prog_bar = False
    model.run()
    before = (model.u.copy(), model.v.copy(), model.w.copy())
    assert (tmp_path / "state_0" / "u.npy").exists()
    model = fw.FentonKarma2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 5
    model.cardiac_tissue = tissue
    model.state_loader = fw.StateLoader(str(tmp_path / "state_0"))
    model.prog_bar = False
    model.run()
    for a, b in zip(before, (model.u, model.v, model.w)):
        assert np.allclose(a, b, atol=1e-5)


@pytest.mark.gpu
def test_current_stimulus_fencepost_matches_oracle(gpu):
    n = 8
    tissue = fw.CardiacTissue2D([n, n])
    model = fw.AlievPanfilov2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 2.0
    model.cardiac_tissue = tissue
    model.prog_bar = False
    seq = fw.StimSequence()
    seq.add_stim(fw.StimCurrentCoord2D(0.3, 0.5, 0.5, 3, 5, 3, 5, u_max=0.2))
    model.stim_sequence = seq
    model.run()
    sim = OracleSim("aliev_panfilov", tissue.mesh, 0.01, 0.25)
    sim.stims.append(OStim("current", 0.3, 0.5, (slice(3, 5), slice(3, 5)), duration=0.5,
                           u_max=0.2))
    sim.run(2.0)
    assert np.array_equal(model.u, sim.u)


@pytest.mark.gpu
def test_literal_entry_points_match_oracle(gpu):
    """fw_compute_weights / fw_diffusion / fw_ionic against the oracle's
    weights and sweeps: weights and diffusion bit for bit."""
    torch = gpu
    from finitewave_amd import _native
    from finitewave_amd.engine import raw_weights
    for dim, stencil, shape in [(2, "iso", (20, 21)), (2, "aniso", (20, 21)),
                                (2, "sym", (20, 21)), (3, "iso", (9, 10, 11)),
                                (3, "aniso", (9, 10, 11))]:
        fib = random_fibers(shape) if stencil != "iso" else None
        model, sim = make_pair("aliev_panfilov", dim, shape, stencil=stencil, fibrosis=0.15,
                               fibers=fib)
        w = model.stencil.compute_weights(model, model.cardiac_tissue)
        assert np.array_equal(w, sim.weights), (dim, stencil)
        # one diffusion sweep with those weights
        un = np.zeros(shape)
        model.u = np.random.RandomState(4).random_sample(shape)
        sim.u = model.u.copy()
        sim.u_new = np.zeros(shape)
        model.stencil.select_diffusion_kernel()(un, model.u, w, sim.idx)
        sim._diffuse()
        assert np.array_equal(un, sim.u_new), (dim, stencil)
    # one ionic sweep of every model
    for name in ["aliev_panfilov", "barkley", "mitchell_schaeffer", "fenton_karma",
                 "bueno_orovio", "luo_rudy91", "tp06", "courtemanche"]:
        model, sim = make_pair(name, 2, (12, 12))
        model.u_new = model.u.copy() * 0.5
        sim.u_new = model.u_new.copy()
        model.run_ionic_kernel()
        sim._ionic()
        assert np.max(np.abs(model.u_new - sim.u_new)) <= 1e-12, name
        for n in model._state_call_order():
            assert np.max(np.abs(getattr(model, n) - sim.state[n])) <= 1e-12 * max(
                1, np.max(np.abs(sim.state[n]))), (name, n)
