"""Batched stepping and the serving loop.

Every tissue of a batch must give the same activation map, probe trace and
final potential as ``model.run()`` on that tissue alone (the single-tissue
path that is itself checked against the oracle), bit for bit, whatever else
shares its launches: tissues of other sizes, fibrosis and stimulus
protocols, admitted at other steps."""
from __future__ import annotations

import copy

import numpy as np
import pytest

import finitewave_amd as fw
from finitewave_amd.analysis import apd90
from finitewave_amd.batch import BatchSim, TissueBatch
from finitewave_amd.serving import Request, Server, fibrotic_fraction, fp64_for_fibrosis
from finitewave_amd.workloads import fibrotic_tissue, restitution_ensemble


def _single(cls, precision, sim: BatchSim, dt=0.01, dr=0.25):
    model = cls()
    model.precision = precision
    model.dt, model.dr, model.t_max = dt, dr, sim.t_max
    model.prog_bar = False
    model.cardiac_tissue = copy.deepcopy(sim.tissue)
    seq = fw.StimSequence()
    for s in sim.stims:
        seq.add_stim(copy.deepcopy(s))
    model.stim_sequence = seq
    ts = fw.TrackerSequence()
    act = fw.ActivationTime2DTracker()
    act.threshold = sim.act_threshold
    act.start_time = sim.act_start
    ts.add_tracker(act)
    ap = fw.ActionPotential2DTracker()
    ap.cell_ind = list(sim.probe)
    ap.step = 1
    ts.add_tracker(ap)
    model.tracker_sequence = ts
    model.run()
    return dict(u=np.array(model.u), act_t=np.array(act.output), trace=np.array(ap.output))


def _sims(kind, n=3, t_max=8.0, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        side = int(rng.integers(20, 60))
        fib = None if kind == "iso" else float(rng.uniform(0, 180))
        tissue = fibrotic_tissue(side, float(rng.uniform(0, 0.25)), 100 + k, fibre_deg=fib)
        n_ = side + 2
        if kind == "tp06":
            thr = -40.0
            stims = [fw.StimCurrentCoord2D(0, 100, 1.0, 0, n_, 0, 5)]
            if k % 2:
                stims.append(fw.StimCurrentCoord2D(4.0, 100, 1.0, n_ // 2, n_, 0, n_ // 2))
        else:
            thr = 0.5
            stims = [fw.StimVoltageCoord2D(0, 1.0, 0, n_, 0, 5)]
            if k % 2:
                stims.append(fw.StimVoltageCoord2D(3.0, 1.0, n_ // 2, n_, 0, n_ // 2))
        mid = n_ // 2
        probe = (mid, int(np.flatnonzero(tissue.mesh[mid] == 1)[0]))
        out.append(BatchSim(tissue=tissue, t_max=t_max + 0.37 * k, stims=stims, probe=probe,
                            act_threshold=thr, act_start=0.5 * k, apd_threshold=thr))
    return out


CASES = [
    ("aliev_panfilov", "fp64", "iso"), ("aliev_panfilov", "fp64", "aniso"),
    ("tp06", "fp64", "tp06"), ("tp06", "fp32", "tp06"),
    ("mitchell_schaeffer", "fp32", "iso"), ("mitchell_schaeffer", "fp32", "aniso"),
]
CLS = {"aliev_panfilov": fw.AlievPanfilov2D, "tp06": fw.TP062D,
       "mitchell_schaeffer": fw.MitchellSchaeffer2D}


def _first_difference(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return f"shapes {a.shape} vs {b.shape}"
    bad = np.argwhere(a != b)
    if len(bad) == 0:
        return "equal"
    k = tuple(bad[0])
    return (f"{len(bad)} of {a.size} differ; first at {k}: {a[k]!r} vs {b[k]!r}; "
            f"max |diff| {np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))):.3e}")


def _check_same(res, ref):
    for key in ("u", "act_t", "trace"):
        a, b = np.asarray(res[key]), np.asarray(ref[key])
        if key == "trace":
            b = b.astype(np.float64)
        assert np.array_equal(a, b), f"{key}: {_first_difference(a, b)}"


@pytest.mark.gpu
@pytest.mark.parametrize("model,prec,kind", CASES, ids=["-".join(c) for c in CASES])
def test_batch_equals_single_tissue_run(gpu, model, prec, kind):
    cls = CLS[model]
    sims = _sims(kind)
    batch = TissueBatch(cls, precision=prec, max_tissues=8, chunk_cap=97)
    # staggered admission: the second and third tissues join a running batch
    batch.admit(sims[0])
    batch.advance(113)
    batch.admit(sims[1])
    batch.advance(58)
    batch.admit(sims[2])
    done = batch.run_all()
    batch.close()
    assert len(done) == 3
    for sim in sims:
        ref = _single(cls, prec, sim)
        _check_same(sim.result, ref)


@pytest.mark.gpu
def test_batch_graph_equals_plain_launches(gpu):
    outs = []
    for graph in (True, False):
        sims = _sims("iso", seed=3)
        b = TissueBatch(fw.AlievPanfilov2D, precision="fp32", use_graph=graph, chunk_cap=64)
        for s in sims:
            b.admit(s)
        b.run_all()
        b.close()
        outs.append([s.result for s in sims])
    for a, b in zip(*outs):
        _check_same(a, b)


@pytest.mark.gpu
def test_serving_result_alone_equals_in_mix(gpu):
    """Each request's result is the same served alone and served among a mix
    of other requests (two models, staggered arrivals, slot reuse)."""
    def reqs():
        out = []
        for k, s in enumerate(_sims("tp06", n=3, t_max=6.0, seed=5)):
            out.append(Request(rid=k, arrival=0.0, model="tp06", sim=s))
        for k, s in enumerate(_sims("aniso", n=4, t_max=7.0, seed=6)):
            out.append(Request(rid=10 + k, arrival=0.0, model="ap", sim=s))
        return out
    classes = {"tp06": fw.TP062D, "ap": fw.AlievPanfilov2D}
    mix = reqs()
    srv = Server(classes, precision="fp32", max_tissues=2, chunk_steps=150,
                 hbm_budget=1 << 30)
    stats = srv.serve(mix, realtime=False)
    srv.close()
    assert stats["completed"] == len(mix)
    assert stats["max_concurrent"] <= 4
    for r in reqs():
        alone = Server(classes, precision="fp32", max_tissues=2, chunk_steps=150,
                       hbm_budget=1 << 30)
        alone.serve([r], realtime=False)
        alone.close()
        other = [m for m in mix if m.rid == r.rid][0]
        _check_same(r.sim.result, other.sim.result)


@pytest.mark.gpu
@pytest.mark.parametrize("fit", [1, 2, 3])
def test_serving_admission_respects_hbm_budget(gpu, fit):
    """Five equal requests under a budget that holds `fit` of them."""
    classes = {"ap": fw.AlievPanfilov2D}
    srv = Server(classes, precision="fp32", max_tissues=8, chunk_steps=50)
    reqs = []
    for k in range(5):
        t = fibrotic_tissue(30, 0.05, 200 + k)
        sim = BatchSim(tissue=t, t_max=2.0, stims=[fw.StimVoltageCoord2D(0, 1.0, 0, 32, 0, 4)],
                       probe=(16, 16), act_threshold=0.5)
        reqs.append(Request(rid=k, arrival=0.0, model="ap", sim=sim))
    need = srv.footprint(reqs[0])
    srv.hbm_budget = (fit + 1) * need - 1
    stats = srv.serve(reqs, realtime=False)
    srv.close()
    assert stats["completed"] == 5
    assert stats["peak_hbm_bytes"] <= srv.hbm_budget
    assert stats["max_concurrent"] == fit


def _ap_request(rid, seed, t_max=3.0, conductivity=1.0):
    t = fibrotic_tissue(30, 0.05, seed)
    t.conductivity = conductivity
    sim = BatchSim(tissue=t, t_max=t_max, stims=[fw.StimVoltageCoord2D(0, 1.0, 0, 32, 0, 4)],
                   probe=(16, 16), act_threshold=0.5)
    return Request(rid=rid, arrival=0.0, model="ap", sim=sim)


@pytest.mark.gpu
def test_serving_cancels_queued_and_running_requests(gpu):
    """A request cancelled while queued is never admitted, one cancelled
    while running leaves its batch at the next launch boundary, one whose
    deadline passes is cancelled; the others' results equal their runs
    alone."""
    classes = {"ap": fw.AlievPanfilov2D}
    reqs = [_ap_request(k, 300 + k) for k in range(5)]
    reqs[3].deadline_s = 0.0            # expires at the first look
    srv = Server(classes, precision="fp32", max_tissues=2, chunk_steps=40,
                 hbm_budget=1 << 30)
    srv.cancel(4)                        # queued behind the two slots
    seen = []

    def log(r):
        seen.append((r.rid, r.status))

    # cancel request 1 once it is running: after the first chunk
    batch = srv.batch(reqs[0])
    orig = batch.advance

    def advance(n=None):
        out = orig(n)
        if batch.step >= 40:
            srv.cancel(1)
        return out
    batch.advance = advance
    stats = srv.serve(reqs, realtime=False, log=log)
    srv.close()
    status = {r.rid: r.status for r in reqs}
    assert status == {0: "done", 1: "cancelled", 2: "done", 3: "cancelled", 4: "cancelled"}
    assert reqs[1].admitted is not None and reqs[4].admitted is None
    assert reqs[1].sim.result is None and reqs[4].sim.result is None
    assert stats["completed"] == 2 and stats["cancelled"] == 3 and stats["diverged"] == 0
    assert sorted(seen) == sorted(status.items())
    for k in (0, 2):
        alone = Server(classes, precision="fp32", max_tissues=2, chunk_steps=40,
                       hbm_budget=1 << 30)
        r = _ap_request(k, 300 + k)
        alone.serve([r], realtime=False)
        alone.close()
        _check_same(reqs[k].sim.result, r.sim.result)


@pytest.mark.gpu
def test_serving_drops_a_diverging_request(gpu):
    """A tissue whose conductivity makes the explicit step unstable
    (D dt / dr^2 far above 1/4) reaches non-finite potentials: the divergence
    check drops it as 'diverged' and the request sharing its batch finishes
    with the results of its run alone."""
    classes = {"ap": fw.AlievPanfilov2D}
    good, bad = _ap_request(0, 400, t_max=4.0), _ap_request(1, 401, t_max=4.0,
                                                            conductivity=200.0)
    srv = Server(classes, precision="fp32", max_tissues=2, chunk_steps=25,
                 hbm_budget=1 << 30)
    stats = srv.serve([good, bad], realtime=False)
    srv.close()
    assert bad.status == "diverged" and "non-finite" in bad.error
    assert bad.sim.result is None
    assert good.status == "done"
    assert stats["completed"] == 1 and stats["diverged"] == 1
    alone = Server(classes, precision="fp32", max_tissues=2, chunk_steps=25,
                   hbm_budget=1 << 30)
    r = _ap_request(0, 400, t_max=4.0)
    alone.serve([r], realtime=False)
    alone.close()
    _check_same(good.sim.result, r.sim.result)


@pytest.mark.gpu
def test_restitution_ensemble_sample(gpu):
    """A few tissues of the restitution ensemble (MS, S1 train + S2): every
    tissue is captured by the S2 beat, APD90 grows with the coupling
    interval, and the S2 activation map of a tissue equals its single run."""
    cls, sims = restitution_ensemble(n=6, sides=(32,), beats=2, bcl=300.0, tail=300.0,
                                     p_range=(0.0, 0.1), ci_range=(330.0, 450.0))
    sims = sorted(sims, key=lambda s: s.tag["ci"])
    b = TissueBatch(cls, precision="fp32", max_tissues=8)
    for s in sims:
        b.admit(s)
    b.run_all()
    b.close()
    apds = [s.result["apd90"] for s in sims]
    assert all(a is not None for a in apds)
    assert apds[-1] > apds[0]
    ref = _single(cls, "fp32", sims[0])
    _check_same(sims[0].result, ref)


@pytest.mark.gpu
def test_restitution_loss_of_capture(gpu):
    """An S2 inside the refractory period of the last S1 beat produces no
    upstroke at the probe: APD90 is None and the tissue is reported as not
    captured (RestitutionCurve.ipynb cell 6), while a late S2 is captured and
    measured on its own beat; the front velocity is found for both S1-S2
    pairs that activate the tissue."""
    # a 500 ms cycle: both S1 beats capture (the model's APD is shorter)
    cls, sims = restitution_ensemble(n=2, sides=(32,), beats=2, bcl=500.0, tail=300.0,
                                     p_range=(0.0, 0.0), ci_range=(30.0, 30.0))
    _, late = restitution_ensemble(n=1, sides=(32,), beats=2, bcl=500.0, tail=400.0,
                                   p_range=(0.0, 0.0), ci_range=(450.0, 450.0))
    b = TissueBatch(cls, precision="fp32", max_tissues=4)
    for s in sims + late:
        b.admit(s)
    b.run_all()
    b.close()
    for s in sims:
        # the last S1 beat was captured at the probe (the whole trace holds
        # an upstroke), the S2 was not
        assert apd90(s.result["trace"], 0.01, 0.5) is not None
        assert s.result["captured"] is False and s.result["apd90"] is None
    r = late[0].result
    assert r["captured"] is True and r["apd90"] is not None and r["apd90"] > 50
    # the S1 beats are in the trace too, but APD90 is the S2 beat's
    assert r["cv_front"] is not None and r["cv_front"] > 0


@pytest.mark.gpu
def test_batch_abi_rejects_updated_nodes_on_the_outer_layer(gpu):
    """The batched kernel has no column guard; fw_batch_set_tissue checks the
    codes of the outer layer on the device and refuses the tissue."""
    import ctypes
    import torch
    from finitewave_amd import _native
    from finitewave_amd.batch import _TissueDesc
    b = TissueBatch(fw.AlievPanfilov2D, precision="fp32", max_tissues=2)
    shape = (8, 9)
    code = np.zeros(shape, np.uint8)
    code[1:-1, 1:-1] = 1
    bufs = [torch.zeros(shape, dtype=torch.float32, device="cuda") for _ in range(3)]
    for bad_node, ok in (((3, 0), False), ((0, 4), False), ((7, 8), False), ((3, 1), True)):
        c = code.copy()
        c[bad_node] = 1
        dc = torch.from_numpy(c).cuda()
        desc = _TissueDesc()
        desc.u, desc.un = bufs[0].data_ptr(), bufs[1].data_ptr()
        desc.state[0] = bufs[2].data_ptr()
        desc.code = dc.data_ptr()
        desc.ni, desc.nj, desc.n = shape[0], shape[1], shape[0] * shape[1]
        desc.stencil = _native.STENCIL_ISO
        desc.probe_idx = -1
        desc.act_start = -1
        rc = _native.lib().fw_batch_set_tissue(b.handle, 0, ctypes.byref(desc), None)
        torch.cuda.synchronize()
        assert (rc == 0) == ok, (bad_node, rc)
        if not ok:
            assert b"outer" in _native.lib().fw_last_error()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("model,prec", [("mitchell_schaeffer", "fp32"), ("aliev_panfilov", "fp64"),
                                        ("tp06", "fp32")])
def test_every_launch_variant_gives_the_same_results(gpu, model, prec):
    cls = CLS[model]
    results = {}
    probe = TissueBatch(cls, precision=prec, max_tissues=2)
    variants = probe.variants()
    probe.close()
    for v in variants:
        sims = _sims("tp06" if model == "tp06" else "aniso", seed=11)
        sims[1].tissue.fibers = None  # iso and aniso tissues in one batch
        b = TissueBatch(cls, precision=prec, max_tissues=4, autotune=False, chunk_cap=90)
        b.set_variant(v)
        for s in sims:
            b.admit(s)
        b.run_all()
        assert b.variant == v
        b.close()
        results[v] = [s.result for s in sims]
    ref = results[variants[0]]
    for v in variants[1:]:
        for a, r in zip(results[v], ref):
            _check_same(a, r)


@pytest.mark.gpu
def test_autotuner_picks_a_variant_and_keeps_results(gpu, tmp_path, monkeypatch):
    from finitewave_amd import batch as B
    monkeypatch.setenv("FW_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(B, "_TUNE_CACHE", {})
    sims = _sims("iso", seed=12, t_max=6.0)
    b = TissueBatch(fw.MitchellSchaeffer2D, precision="fp32", max_tissues=4, tune_steps=8)
    for s in sims:
        b.admit(s)
    b.run_all()
    assert b.tuned and b.tuning is not None and set(b.tuning) == set(b.variants())
    t = b.tuning
    best = min(t, key=t.get)
    # the fastest shape, or an XCD-contiguous one within 3 % of it
    assert b.variant == best or (b.variant & 2 and t[b.variant] <= 1.03 * t[best])
    assert (tmp_path / "tune.json").exists()
    b.close()
    for s in sims:
        _check_same(s.result, _single(fw.MitchellSchaeffer2D, "fp32", s))


@pytest.mark.gpu
def test_autotuner_chooses_again_when_the_batch_grows(gpu, tmp_path, monkeypatch):
    """The launch shape is chosen per size class (power of two of the node
    count): a batch that grows from one small tissue to many chooses again,
    and the results stay those of the single-tissue runs."""
    from finitewave_amd import batch as B
    monkeypatch.setenv("FW_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(B, "_TUNE_CACHE", {})
    sims = _sims("iso", n=6, seed=13, t_max=5.0)
    b = TissueBatch(fw.MitchellSchaeffer2D, precision="fp32", max_tissues=8, tune_steps=4)
    b.admit(sims[0])
    while b.step < 200:
        b.advance(200 - b.step)
    first = b._tuned_key
    assert b.tuned and first is not None
    for s in sims[1:]:
        b.admit(s)
    b.advance()
    assert b._tuned_key != first  # chosen again for the grown batch
    b.run_all()  # shrinking back as tissues finish selects from the cache again
    assert len(B._TUNE_CACHE) >= 2
    b.close()
    for s in sims:
        _check_same(s.result, _single(fw.MitchellSchaeffer2D, "fp32", s))


def _wide_sims(kind, seed):
    """Tissues wider than one 64-column strip and taller than one run of
    rows, so that the multi-step shapes' halos and run boundaries fall
    inside them; second stimuli at steps that split the pairs of launches."""
    rng = np.random.default_rng(seed)
    shapes = [(150, 93), (71, 200), (64, 66), (129, 130)]
    out = []
    for k, (ni, nj) in enumerate(shapes):
        t = fw.CardiacTissue2D([ni, nj])
        mesh = (1 + (rng.random((ni, nj)) <= 0.15)).astype(np.int8)
        mesh[0], mesh[-1], mesh[:, 0], mesh[:, -1] = 0, 0, 0, 0
        t.mesh = mesh
        if kind == "aniso" and k % 2 == 0:
            a = np.deg2rad(25.0 + 40 * k)
            f = np.zeros((ni, nj, 2))
            f[..., 0], f[..., 1] = np.cos(a), np.sin(a)
            t.fibers = f
        stims = [fw.StimVoltageCoord2D(0, 1.0, 0, ni, 0, 5)]
        stims.append(fw.StimVoltageCoord2D(0.37 + 0.11 * k, 1.0, ni // 2, ni, 0, nj // 2))
        probe = (ni // 2, int(np.flatnonzero(mesh[ni // 2] == 1)[-1]))
        out.append(BatchSim(tissue=t, t_max=4.0 + 0.53 * k, stims=stims, probe=probe,
                            act_threshold=0.5, act_start=0.21 * k, apd_threshold=0.5))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("model,prec,kind", [("mitchell_schaeffer", "fp32", "iso"),
                                             ("mitchell_schaeffer", "fp32", "aniso"),
                                             ("aliev_panfilov", "fp64", "aniso"),
                                             ("bueno_orovio", "fp32", "aniso")])
def test_multi_step_shapes_on_wide_tissues(gpu, model, prec, kind):
    """Every multi-step launch shape (K steps per launch, tissues split into
    64-column strips and runs of rows with K-node halos) gives the single-step
    shape's results bit for bit on tissues spanning several strips and runs,
    admitted at different steps, with stimuli between pairs of launches."""
    cls = {"mitchell_schaeffer": fw.MitchellSchaeffer2D, "aliev_panfilov": fw.AlievPanfilov2D,
           "bueno_orovio": fw.BuenoOrovio2D}[model]
    probe = TissueBatch(cls, precision=prec, max_tissues=2)
    variants = [v for v in probe.variants() if v & 8]
    probe.close()
    assert variants
    results = {}
    for v in [0] + variants:
        sims = _wide_sims(kind, seed=5)
        b = TissueBatch(cls, precision=prec, max_tissues=4, autotune=False, chunk_cap=77)
        b.set_variant(v)
        b.admit(sims[0])
        b.admit(sims[1])
        b.advance(45)
        b.admit(sims[2])
        b.advance(19)
        b.admit(sims[3])
        b.run_all()
        assert b.variant == v
        b.close()
        results[v] = [s.result for s in sims]
        print(f"variant {v}: done", flush=True)
    ref = results[0]
    assert any(np.any(r["act_t"] >= 0) for r in ref)
    bad = []
    for v in variants:
        for k, (a, r) in enumerate(zip(results[v], ref)):
            try:
                _check_same(a, r)
            except AssertionError as e:
                bad.append(f"variant {v} tissue {k}: {e}")
    assert not bad, "\n".join(bad)


@pytest.mark.gpu
def test_serving_policy_runs_fibrotic_tp06_tissues_on_the_fp64_path(gpu):
    """fp64_for_fibrosis routes the TP06 tissues at or above the threshold to
    an FP64 batch and every other request to FP32; the two TP06 batches run
    side by side, and each request's results equal its plain single-tissue
    run in the precision it was routed to (for the FP64-routed tissue: the
    FP64 path the served results are compared with)."""
    def tp06(rid, p, seed):
        t = fibrotic_tissue(40, p, seed)
        sim = BatchSim(tissue=t, t_max=6.0, stims=[fw.StimCurrentCoord2D(0, 100, 1.0, 0, 42, 0, 5)],
                       probe=(21, int(np.flatnonzero(t.mesh[21] == 1)[0])), act_threshold=-40.0,
                       apd_threshold=-40.0)
        return Request(rid=rid, arrival=0.0, model="tp06", sim=sim)

    reqs = [tp06(0, 0.05, 500), tp06(1, 0.2, 501), tp06(2, 0.35, 502), _ap_request(3, 503)]
    reqs[3].sim.tissue = fibrotic_tissue(30, 0.35, 503)
    classes = {"tp06": fw.TP062D, "ap": fw.AlievPanfilov2D}
    srv = Server(classes, precision="fp32", max_tissues=4, chunk_steps=120, hbm_budget=1 << 30,
                 precision_for=fp64_for_fibrosis(("tp06",), 0.3))
    stats = srv.serve(reqs, realtime=False)
    srv.close()
    assert stats["completed"] == 4
    assert [r.precision for r in reqs] == ["fp32", "fp32", "fp64", "fp32"]
    assert fibrotic_fraction(reqs[2].sim) >= 0.3 > fibrotic_fraction(reqs[1].sim)
    assert set(srv.batches) == {("tp06", "fp32"), ("tp06", "fp64"), ("ap", "fp32")}
    for r in reqs:
        _check_same(r.sim.result, _single(classes[r.model], r.precision, r.sim))
