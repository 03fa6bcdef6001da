"""Accuracy of the FP32 path against the FP64 path.

The FP64 path is the oracle-pinned reference arithmetic (per-step |dV| <=
1e-12 against the oracle, tests/test_gpu_parity.py), so it serves as the
FP64 reference here for runs too long for the CPU oracle.  Tolerances stated
activation-time error <= 1 dt, APD90 error
<= 1 ms; the short-horizon V error is reported and bounded loosely.
"""
from __future__ import annotations

import numpy as np
import pytest

import finitewave_amd as fw
from tests.scenarios import boxed_mesh, layered_fibers, random_fibers


def _apd90(u, dt, threshold_up):
    # APD90 as in the reference's restitution tutorial (Tutorials/RestitutionCurve.ipynb):
    # last upstroke, 90 % repolarisation level between peak and minimum after it
    up = np.where((u[:-1] < threshold_up) & (u[1:] >= threshold_up))[0]
    if len(up) == 0:
        return None
    s = up[-1]
    peak = np.max(u[s:])
    level = peak - 0.9 * (peak - np.min(u[s:]))
    idx = np.where(u[s:] < level)[0]
    return None if len(idx) == 0 else idx[0] * dt


def _planar(cls, precision, shape, t_max, cur, dur, thr, cell, fibers=None):
    tissue = fw.CardiacTissue2D(list(shape))
    tissue.fibers = fibers
    model = cls()
    model.dt, model.dr, model.t_max = 0.01, 0.25, t_max
    model.precision = precision
    model.cardiac_tissue = tissue
    model.prog_bar = False
    seq = fw.StimSequence()
    seq.add_stim(fw.StimCurrentCoord2D(0, cur, dur, 0, shape[0], 0, 6))
    model.stim_sequence = seq
    act = fw.ActivationTime2DTracker()
    act.threshold = thr
    ap = fw.ActionPotential2DTracker()
    ap.cell_ind = cell
    ts = fw.TrackerSequence()
    ts.add_tracker(act)
    ts.add_tracker(ap)
    model.tracker_sequence = ts
    model.run()
    return model, act.output, ap.output


@pytest.mark.gpu
@pytest.mark.parametrize("aniso", [False, True])
def test_tp06_fp32_activation_and_apd90(gpu, aniso):
    shape = (40, 200)
    fibers = None
    if aniso:
        fibers = np.zeros((*shape, 2))
        fibers[..., 0] = np.cos(np.deg2rad(30))
        fibers[..., 1] = np.sin(np.deg2rad(30))
    out = {}
    for prec in ("fp64", "fp32"):
        out[prec] = _planar(fw.TP062D, prec, shape, 400.0, 100, 1.0, -40, [20, 150], fibers)
    m64, act64, ap64 = out["fp64"]
    m32, act32, ap32 = out["fp32"]
    assert np.isfinite(m32.u).all()
    hit = act64 >= 0
    assert hit.sum() > 0.9 * (shape[0] - 2) * (shape[1] - 2)
    assert np.array_equal(hit, act32 >= 0)
    err_at = float(np.max(np.abs(act64[hit] - act32[hit])))
    assert err_at <= 0.01 + 1e-9, f"activation-time error {err_at} ms > 1 dt"
    a64, a32 = _apd90(ap64, 0.01, -40), _apd90(ap32, 0.01, -40)
    assert a64 is not None and abs(a64 - a32) <= 1.0, (a64, a32)


@pytest.mark.gpu
@pytest.mark.parametrize("cls,thr", [(fw.AlievPanfilov2D, 0.5), (fw.MitchellSchaeffer2D, 0.5)])
def test_fp32_two_variable_models(gpu, cls, thr):
    shape = (30, 160)
    t_max = 60.0 if cls is fw.AlievPanfilov2D else 400.0
    out = {p: _planar(cls, p, shape, t_max, 5, 0.5, thr, [15, 120]) for p in ("fp64", "fp32")}
    act64, act32 = out["fp64"][1], out["fp32"][1]
    hit = act64 >= 0
    assert hit.sum() > 0
    assert np.array_equal(hit, act32 >= 0)
    assert float(np.max(np.abs(act64[hit] - act32[hit]))) <= 0.01 + 1e-9
    a64, a32 = _apd90(out["fp64"][2], 0.01, thr), _apd90(out["fp32"][2], 0.01, thr)
    assert a64 is not None and abs(a64 - a32) <= 1.0


@pytest.mark.gpu
@pytest.mark.parametrize("cls", [fw.LuoRudy912D, fw.Courtemanche2D, fw.FentonKarma2D,
                                 fw.BuenoOrovio2D, fw.Barkley2D])
def test_fp32_other_models_stay_close(gpu, cls):
    # short horizon (5 ms) V error after an upstroke, reported as a bound
    shape = (24, 60)
    phys = cls in (fw.LuoRudy912D, fw.Courtemanche2D)
    cur, thr = (100, -40) if phys else (5, 0.5)
    # stimulus durations of the reference's model tests (tests/test_models_2d.py)
    dur = 1.0 if phys else (0.1 if cls is fw.Barkley2D else 0.5)
    out = {p: _planar(cls, p, shape, 5.0, cur, dur, thr, [12, 30]) for p in ("fp64", "fp32")}
    u64, u32 = out["fp64"][0].u, out["fp32"][0].u
    assert np.isfinite(u64).all() and np.isfinite(u32).all()
    scale = max(1.0, float(np.max(np.abs(u64))))
    assert float(np.max(np.abs(u64 - u32))) <= 2e-3 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fp32_one_step_from_random_state_matches_fp64_oracle(gpu, seed):
    """One FP32 fused step (9-point flux form + Barkley update) from random
    potentials on a fibrotic mesh with a random fibre angle, against the
    oracle's FP64 step: the FP32 path's stencil algebra, masks and minor
    fluxes agree with the reference weights to float rounding.
    Tolerance: 1e-5 absolute (about 40 float ulps of the stepped values, which
    stay within [-2, 2]); a wrong weight or mask is off by 1e-3 or more."""
    from finitewave_amd.engine import Engine
    from tests.scenarios import make_pair
    rng = np.random.RandomState(seed)
    shape = (30, 33)
    ang = rng.uniform(0, np.pi)
    f = np.zeros((*shape, 2))
    f[..., 0], f[..., 1] = np.cos(ang), np.sin(ang)
    model, sim = make_pair("barkley", 2, shape, stencil="aniso", fibrosis=0.3, fibers=f,
                           precision="fp32")
    u0 = rng.uniform(-1, 1, size=shape)
    u0[model.cardiac_tissue.mesh != 1] = 0.0
    model.u[...] = u0
    sim.u[...] = u0
    eng = Engine(model)
    eng.upload()
    eng.run_steps(0, 1, use_graph=False)
    u_gpu = eng.u_dev().cpu().numpy().astype(np.float64)
    sim.one_step()
    sim.u, sim.u_new = sim.u_new, sim.u
    act = model.cardiac_tissue.mesh == 1
    assert float(np.max(np.abs(u_gpu[act] - sim.u[act]))) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("shape,layout", [((20, 17, 13), "layered"), ((35, 9, 40), "layered"),
                                          ((5, 70, 6), "uniform"), ((18, 33, 33), "uniform"),
                                          ((34, 5, 3), "layered")])
def test_fp32_3d_march_one_step_matches_fp64_oracle(gpu, shape, layout):
    """The 3D FP32 19-point step that computes each face flux and square once
    (k_step3_march: 64-node tiles of the (j, k) plane marching 16-plane chunks
    along i) against the oracle's FP64 step, from random potentials on a 30 %
    fibrotic mesh: shapes with several chunks along i, tiles that straddle
    rows and planes, and a plane of three nodes across.  Tolerance 1e-5
    absolute, as for the 2D flux form."""
    from finitewave_amd.engine import Engine
    from tests.scenarios import make_pair
    rng = np.random.RandomState(sum(shape))
    if layout == "layered":
        f = layered_fibers(shape, -70.0, 50.0)
    else:
        ang = rng.uniform(0, np.pi)
        f = np.zeros((*shape, 3))
        f[..., 0], f[..., 1], f[..., 2] = np.cos(ang) * 0.8, np.sin(ang) * 0.8, 0.6
    model, sim = make_pair("aliev_panfilov", 3, shape, stencil="aniso", fibrosis=0.3,
                           fibers=f, precision="fp32")
    u0 = rng.uniform(-1, 1, size=shape)
    u0[model.cardiac_tissue.mesh != 1] = 0.0
    model.u[...] = u0
    sim.u[...] = u0
    eng = Engine(model)
    eng.upload()
    eng.run_steps(0, 1, use_graph=False)
    u_gpu = eng.u_dev().cpu().numpy().astype(np.float64)
    sim.one_step()
    sim.u, sim.u_new = sim.u_new, sim.u
    act = model.cardiac_tissue.mesh == 1
    assert float(np.max(np.abs(u_gpu[act] - sim.u[act]))) <= 1e-5
    # nodes that are not updated keep the u_new buffer's contents
    assert np.all(u_gpu[~act] == 0.0)
    eng.close()


@pytest.mark.gpu
def test_fp32_3d_march_activation_matches_fp64(gpu):
    """A planar wave through a fibrotic slab with layered fibres, 1000 steps
    (the wave crosses the 24 planes in about 600): the FP32 march kernel's
    activation map within one step of the FP64 path's (exact to the
    reference), over every activated node."""
    maps = {}
    shape = (24, 20, 22)
    for precision in ("fp64", "fp32"):
        tissue = fw.CardiacTissue3D(list(shape))
        tissue.mesh = boxed_mesh(shape, 0.15, seed=3)
        tissue.fibers = layered_fibers(shape)
        m = fw.AlievPanfilov3D()
        m.dt, m.dr, m.t_max, m.prog_bar = 0.01, 0.25, 10.0, False
        m.precision = precision
        m.cardiac_tissue = tissue
        seq = fw.StimSequence()
        seq.add_stim(fw.StimVoltageCoord3D(0, 1.0, 0, 3, 0, shape[1], 0, shape[2]))
        m.stim_sequence = seq
        tr = fw.ActivationTime3DTracker()
        tr.threshold = 0.5
        ts = fw.TrackerSequence()
        ts.add_tracker(tr)
        m.tracker_sequence = ts
        m.run()
        maps[precision] = np.asarray(tr.output)
    a, b = maps["fp64"], maps["fp32"]
    n_tissue = np.count_nonzero(boxed_mesh(shape, 0.15, seed=3) == 1)
    assert np.count_nonzero(a >= 0) > 0.9 * n_tissue
    assert np.array_equal(a >= 0, b >= 0)
    both = (a >= 0) & (b >= 0)
    assert np.max(np.abs(a[both] - b[both])) <= 0.01 + 1e-9


def _layer_profile(n, lo, hi):
    return lo + (hi - lo) * np.linspace(0.0, 1.0, n)


FIELD_CASES = {
    # name: (dim, shape, stencil, builder of (fibers, conductivity), expected field modes)
    "3d-aniso-fibres-along-k": (3, (12, 13, 14), "aniso",
                                lambda s: (layered_fibers(s), 1.0), {1}),
    "3d-aniso-fibres-per-node": (3, (11, 12, 13), "aniso",
                                 lambda s: (random_fibers(s, seed=4), 1.0), {2}),
    "3d-iso-conductivity-along-k": (3, (12, 11, 15), "iso",
                                    lambda s: (None, np.broadcast_to(
                                        _layer_profile(s[2], 0.4, 1.0), s).copy()), {1}),
    "2d-aniso-fibres-along-i": (2, (31, 29), "aniso",
                                lambda s: (_fibres_2d(np.linspace(0, np.pi, s[0])[:, None],
                                                      s), 1.0), {0, 1}),
    "2d-aniso-fibres-along-j": (2, (27, 33), "aniso",
                                lambda s: (_fibres_2d(np.linspace(0, np.pi, s[1])[None, :],
                                                      s), 1.0), {0, 1}),
    "2d-iso-conductivity-along-j": (2, (25, 30), "iso",
                                    lambda s: (None, np.broadcast_to(
                                        _layer_profile(s[1], 0.3, 1.0)[None, :], s).copy()),
                                    {0, 1}),
}


def _fibres_2d(ang, shape):
    f = np.zeros((*shape, 2))
    f[..., 0] = np.cos(ang)
    f[..., 1] = np.sin(ang)
    return f


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(FIELD_CASES))
def test_fp32_one_step_with_tabulated_fields_matches_fp64_oracle(gpu, case):
    """One FP32 fused step with half-step fields stored as a table along one
    axis or per node (the plan's float copies, and one load where a table's
    axis differs from the face direction), 2D and 3D, iso and aniso, from
    random potentials on a fibrotic mesh, against the oracle's FP64 step.
    Tolerance 1e-5 absolute, as for uniform fields; a field read at the wrong
    node or face is off by 1e-3 or more."""
    from finitewave_amd.engine import Engine
    from finitewave_amd.stencil import compress_field
    from tests.scenarios import make_pair
    dim, shape, stencil, build, modes = FIELD_CASES[case]
    fibers, cond = build(shape)
    model, sim = make_pair("barkley", dim, shape, stencil=stencil, fibrosis=0.2, fibers=fibers,
                           conductivity=cond, precision="fp32")
    # the case exercises the field storage it names (uniform = 0, table = 1, full = 2)
    got = {compress_field(a)[0] for a in model.stencil.half_step_fields(model,
                                                                        model.cardiac_tissue)}
    assert got & {1, 2} and got <= modes | {0}, got
    rng = np.random.RandomState(len(case))
    u0 = rng.uniform(-1, 1, size=shape)
    u0[model.cardiac_tissue.mesh != 1] = 0.0
    model.u[...] = u0
    sim.u[...] = u0
    eng = Engine(model)
    eng.upload()
    eng.run_steps(0, 1, use_graph=False)
    u_gpu = eng.u_dev().cpu().numpy().astype(np.float64)
    sim.one_step()
    sim.u, sim.u_new = sim.u_new, sim.u
    act = model.cardiac_tissue.mesh == 1
    assert float(np.max(np.abs(u_gpu[act] - sim.u[act]))) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name,fibres,precision", [
    ("TP062D", "uniform", "fp32"), ("AlievPanfilov2D", "uniform", "fp32"),
    ("AlievPanfilov2D", None, "fp32"), ("TP062D", "uniform", "fp64"),
    ("AlievPanfilov3D", "uniform", "fp32"), ("AlievPanfilov3D", "layered", "fp32"),
    ("AlievPanfilov3D", "layered", "fp64"), ("AlievPanfilov3D", None, "fp64"),
    ("AlievPanfilov3D", "uniform", "fp64"), ("AlievPanfilov2D", "uniform", "fp64")])
def test_field_layout_kernels_equal_general_kernel(gpu, name, fibres, precision):
    """The step kernels instantiated for a field layout (every field uniform;
    in 3D, every field uniform or a table along k, as with layered fibres)
    give the general kernel's bits (FP64: the uniform instance selects the
    cross-term quotients the host divided out).  The second tissue's conductivity
    differs at the corner node (0, 0(, 0)) only: no updated node's half-step
    fields read it, so the potentials are the same, but its plan falls back
    to the general kernel, which reads its fields by per-field dispatch."""
    from finitewave_amd.stencil import compress_field
    dim = 3 if name.endswith("3D") else 2
    shape = (48, 40) if dim == 2 else (14, 12, 16)
    runs = []
    for corner in (1.0, 0.5):
        tissue = (fw.CardiacTissue2D if dim == 2 else fw.CardiacTissue3D)(list(shape))
        if fibres == "uniform":
            f = np.zeros((*shape, dim))
            f[..., 0], f[..., 1] = np.cos(np.deg2rad(30.0)), np.sin(np.deg2rad(30.0))
            tissue.fibers = f
        elif fibres == "layered":
            tissue.fibers = layered_fibers(shape)
        cond = np.ones(shape)
        cond[(0,) * dim] = corner
        tissue.conductivity = cond
        model = getattr(fw, name)()
        model.dt, model.dr, model.t_max = 0.01, 0.25, 3.0
        model.precision = precision
        model.prog_bar = False
        model.cardiac_tissue = tissue
        seq = fw.StimSequence()
        box = (0, shape[0], 0, 5) + ((0, shape[2]) if dim == 3 else ())
        stim = fw.StimCurrentCoord2D if dim == 2 else fw.StimCurrentCoord3D
        seq.add_stim(stim(0, 100 if name.startswith("TP06") else 5, 0.5, *box))
        model.stim_sequence = seq
        model.initialize()
        modes = {compress_field(a)[0]
                 for a in model.stencil.half_step_fields(model, tissue)}
        if corner == 1.0:   # a specialised kernel runs
            assert modes <= ({0} if fibres != "layered" else {0, 1}), modes
        else:               # the general one
            assert 2 in modes, modes
        model.run(initialize=False)
        runs.append((model.u.copy(), {n: getattr(model, n).copy()
                                      for n in model._state_call_order()}))
    (u_a, st_a), (u_b, st_b) = runs
    assert np.any(u_a != u_a.flat[0])
    if dim == 3 and fibres is not None and precision == "fp32":
        # the 3D FP32 19-point step for these layouts computes each face and
        # square once (k_step3_march): the same operator, sums in another
        # order; 300 steps of a propagating front stay within float rounding
        assert np.max(np.abs(u_a - u_b)) <= 1e-4, np.max(np.abs(u_a - u_b))
        for n in st_a:
            assert np.max(np.abs(st_a[n] - st_b[n])) <= 1e-4, n
        return
    assert np.array_equal(u_a, u_b)
    for n in st_a:
        assert np.array_equal(st_a[n], st_b[n]), n


def _tp06_run(precision, t_stops, track_vars=None, cell=(10, 10)):
    """A 24x24 TP06 tissue stimulated along one edge, run to each time in
    t_stops in turn (the first call initializes, the others continue)."""
    shape = (24, 24)
    model = fw.TP062D()
    model.dt, model.dr = 0.01, 0.25
    model.precision = precision
    model.cardiac_tissue = fw.CardiacTissue2D(list(shape))
    model.prog_bar = False
    seq = fw.StimSequence()
    seq.add_stim(fw.StimCurrentCoord2D(0, 52, 1.0, 0, shape[0], 0, 4))
    model.stim_sequence = seq
    tracker = None
    if track_vars:
        tracker = fw.MultiVariable2DTracker()
        tracker.cell_ind = list(cell)
        tracker.var_list = list(track_vars)
        ts = fw.TrackerSequence()
        ts.add_tracker(tracker)
        model.tracker_sequence = ts
    for n, t in enumerate(t_stops):
        model.t_max = t
        model.run(initialize=(n == 0))
    return model, tracker


@pytest.mark.gpu
def test_fp32_tp06_stored_differences_read_back_as_values(gpu):
    """The FP32 path keeps TP06's nai, Ki, h and j on the device as float
    differences from their initial values (fw_state_offset).  Every host view
    holds values: the attributes after run(), the device probes of a variable
    tracker; a run continued with run(initialize=False) re-encodes them and
    follows the uninterrupted run; and the values stay within FP32 accuracy of
    the FP64 path.  A missing or doubled offset moves nai by 7.67, Ki by 138.3
    and h, j by 0.75."""
    names = ("nai", "Ki", "h", "j")
    cell = (10, 10)
    whole, tracker = _tp06_run("fp32", [6.0], track_vars=names, cell=cell)
    split, _ = _tp06_run("fp32", [2.0, 6.0])
    ref, _ = _tp06_run("fp64", [6.0])
    assert np.max(whole.u[cell]) > 0.0  # the wave has reached the probed node
    for n in names:
        a, b, r = getattr(whole, n), getattr(split, n), getattr(ref, n)
        assert a.dtype == np.float64, n
        # re-encoding a value rounds only differences far below a float's
        # resolution of the value (|d| < 2^-22 of the offset)
        assert np.max(np.abs(a - b)) <= 1e-9 * max(1.0, float(np.max(np.abs(r)))), n
        # at the front a timing difference well inside one step moves the
        # fast sodium gates by up to a few hundredths
        tol = {"nai": 1e-4, "Ki": 1e-4, "h": 0.05, "j": 0.05}[n]
        assert np.max(np.abs(a - r)) <= tol, (n, float(np.max(np.abs(a - r))))
        # the probe records the value the attribute ends with
        assert tracker.output[n][-1] == a[cell], n
    assert np.allclose(whole.u, split.u, rtol=0, atol=1e-9)
