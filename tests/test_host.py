"""Host-side logic without a GPU: the C-ABI library's symbol table, tissue
codes, half-step fields, field compression, tracker/saver gates, stimulus
boxes and the failure mode when no device is present."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

import finitewave_amd as fw
from finitewave_amd import _native
from finitewave_amd.model import _gate_open, _saver_due
from finitewave_amd.stencil import compress_field
from tests.scenarios import boxed_mesh, layered_fibers, random_fibers

ROOT = Path(__file__).resolve().parent.parent


def test_library_loads_and_exports_every_declared_symbol():
    hdr = (ROOT / "include" / "finitewave_amd.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(fw_[a-z0-9_]+)\s*\(", hdr, re.M))
    assert declared == set(_native.EXPORTS)
    L = _native.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.fw_abi_version() == _native.ABI_VERSION
    assert _native.model_info(_native.MODEL_IDS["tp06"]) == (19, 49)
    assert _native.model_info(_native.MODEL_IDS["courtemanche"]) == (21, 39)
    # TP06's calcium concentrations cai/casr/cass stay FP64 on the FP32 path;
    # nai, Ki and the sodium gates h, j are float differences from 7.67,
    # 138.3, 0.75, 0.75 (the initial values)
    assert [_native.state_is_f64(6, 1, k) for k in range(9)] == [True] * 3 + [False] * 6
    assert all(_native.state_is_f64(6, 0, k) for k in range(19))
    assert [_native.state_offset(6, 1, k) for k in range(9)] == \
        [0.0, 0.0, 0.0, 7.67, 138.3, 0.0, 0.75, 0.75, 0.0]
    assert all(_native.state_offset(m, p, k) == 0.0 for m in range(8) for p in (0, 1)
               for k in range(_native.model_info(m)[0]) if (m, p) != (6, 1))
    # the potential: FP64 on the FP64 path and for LR91/TP06/CRN on the FP32 path
    assert [_native.u_is_f64(m, 1) for m in range(8)] == [False] * 5 + [True] * 3
    assert all(_native.u_is_f64(m, 0) for m in range(8))


def test_abi_rejects_bad_arguments_without_a_device():
    L = _native.lib()
    desc = _native.fw_plan_desc()
    desc.ndim = 4
    import ctypes
    h = ctypes.c_void_p()
    assert L.fw_plan_create(ctypes.byref(desc), ctypes.byref(h)) == -1
    assert b"ndim" in L.fw_last_error()
    assert L.fw_model_info(99, None, None) == -1


def test_tissue_boundaries_and_codes():
    t = fw.CardiacTissue2D([6, 7])
    assert t.mesh[0].sum() == 0 and t.mesh[:, -1].sum() == 0
    t.mesh[2, 3] = 2
    special = np.zeros((6, 7), dtype=np.int8)
    special[3, 3] = 1
    t.special_boundaries = special
    code = t.node_codes()
    assert code[2, 3] == 0 and code[3, 3] == 3 and code[1, 1] == 1
    t.compute_myo_indexes()
    assert np.ravel_multi_index((3, 3), (6, 7)) not in t.myo_indexes
    t.mesh[0, 2] = 1  # in-place edit past the setter
    with pytest.raises(ValueError):
        t.check_boundary()
    t3 = fw.CardiacTissue3D([4, 5, 6])
    assert t3.mesh.sum() == 2 * 3 * 4


def test_half_step_fields_match_oracle_bitwise():
    from oracle.oracle import half_step_fields
    for shape, fib in [((9, 10), random_fibers((9, 10))),
                       ((6, 7, 8), layered_fibers((6, 7, 8)))]:
        t = (fw.CardiacTissue2D if len(shape) == 2 else fw.CardiacTissue3D)(list(shape))
        t.fibers = fib
        t.conductivity = 0.5 + np.random.RandomState(0).random_sample(shape)
        st = fw.AsymmetricStencil2D() if len(shape) == 2 else fw.AsymmetricStencil3D()
        m = fw.AlievPanfilov2D()
        mine = st.half_step_fields(m, t)
        ref = half_step_fields(1, 1 / 9, fib, t.conductivity, shape, len(shape))
        assert len(mine) == len(ref)
        for a, b in zip(mine, ref):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_compress_field_modes():
    assert compress_field(np.full((4, 5), 0.3))[:3] == (_native.FIELD_UNIFORM, 0, 0.3)
    a = np.tile(np.arange(6.0), (4, 5, 1))
    mode, axis, _, data = compress_field(a)
    assert (mode, axis) == (_native.FIELD_AXIS, 2) and np.array_equal(data, np.arange(6.0))
    b = np.random.RandomState(1).random_sample((4, 5))
    assert compress_field(b)[0] == _native.FIELD_FULL
    # signed zeros are different bit patterns: no silent merge
    z = np.zeros((3, 3))
    z[1, 1] = -0.0
    assert compress_field(z)[0] == _native.FIELD_FULL


def test_tracker_and_saver_gates():
    tr = fw.ActivationTime2DTracker()
    tr.start_time, tr.end_time, tr.step = 1.0, 2.0, 3
    assert not _gate_open(tr, 0.5, 3)
    assert _gate_open(tr, 1.5, 6) and not _gate_open(tr, 1.5, 7)
    assert not _gate_open(tr, 2.5, 6)

    class M:
        t, t_max = 4.0, 5.0
    s = fw.StateSaver("x", time=3)
    s.passed = False
    assert _saver_due(s, M)
    s2 = fw.StateSaver("x", time=-1)
    assert not _saver_due(s2, M)
    col = fw.StateSaverCollection()
    col.savers = [s2, s]
    assert _saver_due(col, M)


def test_stim_box_descriptors_clip_like_numpy_slices():
    from finitewave_amd.stimulation import _box_desc
    d = _box_desc(1, (slice(0, 2), slice(-3, None)), (5, 7), 0.5, u_max=1.0)
    assert list(d.lo[:3]) == [0, 4, 0] and list(d.hi[:3]) == [2, 7, 1]
    assert d.clamp == 1 and d.value == 0.5


def test_state_saver_and_loader_roundtrip(tmp_path):
    class M:
        state_vars = ["u", "v"]
        t, t_max = 1.0, 1.0
    m = M()
    m.u = np.arange(6.0).reshape(2, 3)
    m.v = -m.u
    s = fw.StateSaver(str(tmp_path / "st"))
    s.initialize(m)
    s.save()
    n = M()
    ld = fw.StateLoader(str(tmp_path / "st"))
    ld.initialize(n)
    ld.load()
    assert np.array_equal(n.u, m.u) and np.array_equal(n.v, m.v)
    with pytest.raises(FileNotFoundError):
        fw.StateLoader(str(tmp_path / "missing")).initialize(n)


def test_run_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    model = fw.AlievPanfilov2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 0.1
    model.cardiac_tissue = fw.CardiacTissue2D([8, 8])
    model.prog_bar = False
    with pytest.raises(fw.NativeUnavailableError):
        model.run()


def test_model_classes_keep_reference_attributes():
    m = fw.TP062D()
    assert m.state_vars[0] == "u" and len(m.state_vars) == 20
    assert m.D_model == 0.154 and m.init_Ki == 138.3 and m.knaca == 1000
    c = fw.Courtemanche3D()
    assert c.Vj == c.Vc * 0.68 and c.state_vars[15:17] == ["xr", "xs"]
    assert isinstance(fw.AlievPanfilov3D().select_stencil(fw.CardiacTissue3D([4, 4, 4])),
                      fw.IsotropicStencil3D)
    t = fw.CardiacTissue2D([4, 4])
    t.fibers = np.zeros((4, 4, 2))
    assert isinstance(fw.Barkley2D().select_stencil(t), fw.AsymmetricStencil2D)


# ---------------------------------------------------------------- batches (host side)
from finitewave_amd.analysis import (apd90, captured, conduction_velocity,  # noqa: E402
                                     front_velocity)
from finitewave_amd.batch import (BatchSim, _uniform_fields, run_length,  # noqa: E402
                                  stim_schedule, time_fold)


@pytest.mark.parametrize("angle", [0.0, 30.0, 97.5])
def test_batch_uniform_fields_equal_stencil_fields_bitwise(angle):
    n = 12
    t = fw.CardiacTissue2D([n, n])
    t.conductivity = 0.7
    a = np.deg2rad(angle)
    f = np.zeros((n, n, 2))
    f[..., 0], f[..., 1] = np.cos(a), np.sin(a)
    t.fibers = f
    fields = fw.AsymmetricStencil2D().half_step_fields(None, t)
    d4 = _uniform_fields(BatchSim(tissue=t, t_max=1.0), _native.STENCIL_ANISO)
    for k in range(4):
        assert np.all(fields[k] == d4[k]) and np.all(np.signbit(fields[k]) == np.signbit(d4[k]))
    t.fibers = None
    iso = fw.IsotropicStencil2D().half_step_fields(None, t)
    d2 = _uniform_fields(BatchSim(tissue=t, t_max=1.0), _native.STENCIL_ISO)
    assert np.all(iso[0] == d2[0]) and np.all(iso[1] == d2[1])


def test_batch_rejects_non_uniform_fields():
    t = fw.CardiacTissue2D([8, 8])
    t.fibers = random_fibers((8, 8))
    with pytest.raises(ValueError):
        _uniform_fields(BatchSim(tissue=t, t_max=1.0), _native.STENCIL_ANISO)
    t.fibers = None
    t.conductivity = np.ones((8, 8))
    with pytest.raises(ValueError):
        _uniform_fields(BatchSim(tissue=t, t_max=1.0), _native.STENCIL_ISO)


def _loop_schedule(stims, dt, n_steps):
    """Stimulus applications of the model loop (stim_sequence.py:74-77):
    the stimuli's own update_status on a stand-in model."""
    class M:
        t = 0.0
    m = M()
    out = []
    for k in range(n_steps):
        for s in stims:
            if m.t >= s.t and not s.passed:
                out.append((k, s))
                s.update_status(m)
        m.t += dt
    return out


def test_batch_stim_schedule_follows_the_model_loop():
    dt = 0.01
    stims = [fw.StimCurrentCoord2D(0, 5, 0.5, 0, 5, 0, 3),
             fw.StimVoltageCoord2D(0.3, 1.0, 0, 4, 0, 4),
             fw.StimCurrentCoord2D(1.234, 2, 0.07, 0, 5, 0, 3),
             fw.StimVoltageCoord2D(9.0, 1.0, 0, 4, 0, 4),
             fw.StimCurrentCoord2D(2.0, 1, 0.0, 0, 5, 0, 3)]
    n = 700
    fold = time_fold(dt, n + 1)
    got = stim_schedule(stims, n, fold)
    for s in stims:
        s.passed = False
    ref = _loop_schedule(stims, dt, n)
    assert [(k, id(s)) for k, s in got] == [(k, id(s)) for k, s in ref]
    # a 0.5 ms current at dt = 0.01 applies 51 times (the reference's fencepost)
    assert sum(1 for _, s in got if s is stims[0]) == 51
    assert not any(s is stims[3] for _, s in got)


@pytest.mark.parametrize("t_max", [1.0, 8.37, 10.0, 3.3333, 0.005])
def test_batch_run_length_equals_model_loop(t_max):
    dt = 0.01
    fold = time_fold(dt, int(np.ceil(t_max / dt)) + 2)
    # the model loop: iters steps, stopping when t > t_max (cardiac_model.py:191-203)
    t, step = 0.0, 0
    for _ in range(int(np.ceil(t_max / dt))):
        t += dt
        step += 1
        if t > t_max or step > int(np.ceil(t_max / dt)):
            break
    assert run_length(t_max, dt, fold) == step


def test_time_fold_is_the_left_fold():
    f = time_fold(0.01, 1001)
    assert f[1000] == 9.999999999999831
    assert f[0] == 0.0 and f[1] == 0.01


def test_apd90_closed_form_on_a_synthetic_trace():
    dt = 0.1
    t = np.arange(0, 600, dt)
    # rest 0, linear upstroke over 100..102 to 1, linear repolarisation to 0 by 402
    u = np.where(t < 100, 0.0, np.minimum((t - 100) / 2, 1 - (t - 102) / 300))
    u = np.clip(u, 0, 1)
    apd = apd90(u, dt, 0.5)
    # from the last sample below 0.5 (t = 100.9) to the first below 0.1 (t ~ 372)
    assert abs(apd - (372.0 - 100.9)) <= 2 * dt
    # the last of two beats is measured
    u2 = np.concatenate([u, 0.5 * u])
    assert abs(apd90(u2, dt, 0.25) - (372.0 - 100.9)) <= 2 * dt
    assert apd90(np.zeros(100), dt, 0.5) is None
    # the tutorial's quirk is kept: a jump from below the 90% level straight
    # over the threshold gives 0
    jump = np.where(t < 100, 0.0, np.clip(1 - (t - 100) / 300, 0, 1))
    assert apd90(jump, dt, 0.5) == 0.0


def test_apd90_of_the_s2_segment_and_loss_of_capture():
    """RestitutionCurve.ipynb cell 6 measures the S2 run's trace only: an S2
    that produces no upstroke is a loss of capture (None), not the APD90 of
    the S1 beat before it."""
    dt = 0.1
    t = np.arange(0, 900, dt)
    beat = np.clip(np.minimum((t - 100) / 2, 1 - (t - 102) / 300), 0, 1) * (t >= 100)
    s2 = int(round(500 / dt))
    # S1 beat at 100 ms, then a sub-threshold bump at 500 ms (S2 not captured)
    bump = 0.3 * np.exp(-np.maximum(t - 500, 0) / 5) * (t >= 500)
    u = beat + bump
    assert apd90(u, dt, 0.5) is not None          # the whole trace holds the S1 beat
    assert apd90(u, dt, 0.5, start=s2) is None    # the S2 segment does not
    assert not captured(u, 0.5, start=s2) and captured(u, 0.5)
    # a captured S2: its own beat is measured
    t2 = t - 400
    beat2 = np.clip(np.minimum((t2 - 100) / 2, 1 - (t2 - 102) / 200), 0, 1) * (t2 >= 100)
    v = np.where(t < 450, beat, beat2)
    assert captured(v, 0.5, start=s2)
    assert abs(apd90(v, dt, 0.5, start=s2) - (102 + 0.9 * 200 - 100.9)) <= 2 * dt


def test_front_velocity_closed_forms():
    """Velocity2DCalculation.front_velocity (velocity_2d_calculation.py:15-44)
    on maps whose answer is known: a planar front, a point source (distance
    from the source to the far corners), nodes that never activated."""
    dr = 0.25
    act = np.tile(np.arange(40) * 0.5, (6, 1)) + 3.0  # 0.5 ms per node, from t = 3
    v = front_velocity(act, dr)
    assert v.shape == (6,) and np.allclose(v, dr / 0.5, rtol=0, atol=1e-12)
    act[:, 0] = -1  # never activated: ignored for the earliest time
    assert np.allclose(front_velocity(act, dr), dr / 0.5)
    i, j = np.meshgrid(np.arange(21), np.arange(21), indexing="ij")
    ring = np.hypot(i - 10, j - 10)
    pt = np.round(ring, 6) * 2.0        # a point source at (10, 10), 2 ms per node
    v = front_velocity(pt, dr)
    assert np.allclose(v, dr / 2.0)     # four corners, each sqrt(200) nodes away
    assert len(v) == 4
    assert front_velocity(-np.ones((4, 4)), dr) is None
    assert front_velocity(np.zeros((4, 4)), dr) is None


def test_conduction_velocity_closed_form():
    dr = 0.25
    act = np.tile(np.arange(41) * 0.5, (3, 1))      # 0.5 ms per node
    assert abs(conduction_velocity(act, 1, dr) - dr / 0.5) < 1e-12
    act[1, 30] = -1
    assert conduction_velocity(act, 1, dr) is None


# ---------------------------------------------------------------- device trackers (host side)

def test_lat_layer_rebuild_equals_the_reference_rule_step_by_step():
    """The device keeps each node's crossing calls; the host rebuilds the
    reference's layers from them.  On random potentials (many crossings,
    re-entries) the rebuilt layers equal the oracle's step-by-step
    restatement of LocalActivationTime2DTracker._track."""
    from finitewave_amd.tracker import _lat_layers
    from oracle.oracle import OTracker, _lat_track
    rng = np.random.RandomState(4)
    shape, steps, thr, cap = (7, 9), 300, 0.5, 300
    u_seq = (rng.random_sample((steps, *shape)) < 0.3).astype(np.float64)
    u_seq += 0.1 * rng.random_sample((steps, *shape))
    times = np.arange(steps) * 0.01
    tr = OTracker("lat", threshold=thr)
    # device bookkeeping (k_lat) emulated per node
    n = int(np.prod(shape))
    state = np.zeros(n, dtype=np.int64)
    st_steps = np.zeros((n, cap), dtype=np.int64)
    for k in range(steps):
        _lat_track(tr, u_seq[k], times[k])
        u = u_seq[k].reshape(-1)
        act = state & 1
        cross = (u >= thr) & (act == 0)
        cnt = state >> 1
        st_steps[np.flatnonzero(cross), cnt[cross]] = k
        state = np.where(cross, ((cnt + 1) << 1) | 1, state)
        state = np.where((u < thr) & (act == 1), state & ~1, state)
    got = _lat_layers(state >> 1, st_steps, times, shape)
    ref = tr.data["act_t"]
    assert len(got) == len(ref) and len(ref) > 5
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_period_tracker_host_rule_equals_oracle():
    """Period2DTracker's host replay over probe samples equals the oracle's
    period tracker on the same samples."""
    from oracle.oracle import OTracker, _lat_track
    rng = np.random.RandomState(8)
    cells = [[1, 2], [3, 4], [5, 1]]
    steps = 200
    u_full = (rng.random_sample((steps, 7, 9)) < 0.25).astype(np.float64)
    t_hist = list(np.arange(steps) * 0.1)

    class M:
        pass
    m = M()
    m._t_hist = []
    tr = fw.Period2DTracker()
    tr.cell_ind = cells
    tr.threshold = 0.5
    tr.initialize(m)  # as model.initialize does, before any step
    m._t_hist = t_hist
    otr = OTracker("period", threshold=0.5, cells=cells)
    cidx = tuple(np.atleast_2d(cells).T)
    for k in range(steps):
        tr._append(u_full[k][cidx])
        _lat_track(otr, u_full[k], t_hist[k])
    for a, b in zip(tr.act_t, otr.data["act_t"]):
        assert np.array_equal(a, b)
    assert len(tr.act_t) == len(otr.data["act_t"])
    periods = tr.output
    assert len(periods) == len(cells)


def test_batch_launch_variants_per_model():
    import ctypes
    L = _native.lib()
    L.fw_batch_variants.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32]
    # light models: the 8 single-step shapes, then the multi-step ones (bit 3)
    # for K = 2, 4 (bits 4-5) and runs of 64, 128 rows (bits 6-7) over the
    # 4-node XCD-contiguous shape with and without the cache hint
    multi = [base | 8 | (k << 4) | (r << 6) for base in (7, 3) for k in (0, 1) for r in (0, 1)]
    # models 5-7: the four one-node shapes, then the same over node lists (bit 8)
    heavy = [0, 2, 4, 6, 256, 258, 260, 262]
    for model, expect in ((2, list(range(8)) + multi), (0, list(range(8)) + multi),
                          (5, heavy), (6, heavy), (7, heavy)):
        arr = (ctypes.c_int32 * 32)()
        n = L.fw_batch_variants(model, arr, 32)
        assert list(arr[:n]) == expect
    arr = (ctypes.c_int32 * 2)()
    assert L.fw_batch_variants(2, arr, 2) == 16 and list(arr) == [0, 1]  # count beyond capacity
    assert L.fw_batch_variants(99, None, 0) == -1


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    why, before any device work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)
    assert r.stdout.strip() == "" or not r.stdout.lstrip().startswith("{")


def test_schedule_windows_follow_the_stimulus_rules():
    """schedule_windows resolves stimulate_next + update_status
    (stim_sequence.py:74-77, stim.py:52-56) ahead of time: the same steps as
    applying the rules step by step with the FP64 time fold, on random
    sequences; and the fenceposts of the time fold: a 0.5 ms current from
    t = 0 at dt 0.01 applies 51 times, an S2 at 310 ms first applies on step
    31,001 (t after 31,000 steps is 309.99999999986306)."""
    from finitewave_amd.stimulation import schedule_windows

    class _M:
        t = 0.0

    rng = np.random.RandomState(0)
    dt = 0.01
    for trial in range(60):
        stims = []
        for _ in range(rng.randint(1, 5)):
            t0 = round(float(rng.uniform(0, 2)), int(rng.randint(0, 4)))
            dur = 0.0 if rng.rand() < 0.3 else round(float(rng.uniform(0, 1)), int(rng.randint(0, 3)))
            stims.append(fw.StimCurrentCoord2D(t0, 1.0, dur, 0, 1, 0, 1))
        t_start = 0.0 if trial % 3 else float(rng.uniform(0, 1))
        iters = int(rng.randint(1, 300))
        got = [(stims.index(s), lo, hi) for s, lo, hi in schedule_windows(stims, t_start, dt, iters)]
        # step by step, the reference's rules on copies
        m = _M()
        t = t_start
        apps = {}
        for q in range(iters):
            m.t = t
            for k, s in enumerate(stims):
                if t >= s.t and not s.passed:
                    apps.setdefault(k, []).append(q)
                    s.update_status(m)
            t = t + dt
        want = [(k, a[0], a[-1]) for k, a in sorted(apps.items())]
        for k, a in apps.items():
            assert a == list(range(a[0], a[-1] + 1))
        assert got == want, trial
    s = fw.StimCurrentCoord2D(0, 1.0, 0.5, 0, 1, 0, 1)
    assert [(lo, hi) for _, lo, hi in schedule_windows([s], 0.0, dt, 1000)] == [(0, 50)]
    s2 = fw.StimCurrentCoord2D(310, 100, 1.0, 0, 1, 0, 1)
    [(_, lo, hi)] = schedule_windows([s2], 0.0, dt, 32000)
    assert lo == 31001
    t = 0.0
    for _ in range(31000):
        t += dt
    assert t == 309.99999999986306
