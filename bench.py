"""Benchmark of the fused stepping kernel on one MI355X.

Workload (the reference's headline configuration that fits one GPU): TP06
epicardial, 1024x1024 tissue plus a one-node zero boundary (1026x1026),
anisotropic 9-point stencil with a uniform fibre angle of 30 degrees
(D_long 0.154, D_trans 0.154/9), dr 0.025 cm, dt 0.01 ms, planar S1
current stimulus on the left 10 tissue columns at t = 0, activation-time
tracker on.  A step is one fused diffusion + ionic + tracker pass over the
whole tissue; inputs are resident in HBM when timing starts.  The S1 current
(steps 0-100) runs untimed before the warm-up steps, so the timed steps are
the propagation steps that make up all but 200 of C2's 100,000.

Units: the reference's TP06 takes D_model = 0.154 in mm^2/ms and dr in mm
(its own example, examples/models/2D/tp06_2d.py:47, sets dr = 0.25), so
dr = 0.025 cm enters the model as dr = 0.25.  Entered as 0.025 the explicit
scheme has D*dt/dr^2 = 2.46 > 1/4 and diverges, in the reference as here.

The JSON line's value is the FP64 path (the reference's arithmetic); the
FP32 path (TP06 calcium and concentrations in FP64) is reported beside it.
roofline: algorithmic bytes per step / measured time per step against the
8 TB/s HBM3E peak.  cpu_baseline: the C/OpenMP oracle on the host cores over
a bounded sample of the same workload.

Beside the headline the line carries the other configurations' legs: the
512-tissue Mitchell-Schaeffer restitution ensemble (`ensemble`), the
Poisson-served TP06/Aliev-Panfilov mix (`serving`, with the same requests
run one after another on the plain FP64 path), the reference's own CPU case
(`c1`: Aliev-Panfilov 102x102 FP64), the headline tissue's S1-S2 spiral
over 100,000 steps in FP64 and FP32 compared by statistics (`c2_spiral`)
and the 3D slab (`c3`: Aliev-Panfilov 258^3, 19-point stencil, layered
fibres, fibrosis, FP32).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
N_SIDE = 1024
S1_END = 100  # last step of the 1 ms S1 current (steps 0-100 by the FP64 time fold)
FIBER_DEG = 30.0
DR = 0.25  # 0.025 cm in the model's millimetres

# algorithmic bytes per active node and step: the state the reference's step
# reads + the state it writes (TP06 reads u and 18 state arrays, writes u_new,
# 17 of them and oo; cai is read-only).  FP64: 38 arrays x 8 B.  Mixed FP32
# form: FP32 state with the calcium concentrations cai, casr, cass in FP64 =
# 172 B.  Plus the 1-byte node code per node and the 4-byte activation word
# per active node.
B_STATE = {"fp64": 304, "fp32": 172}
# bytes the FP32 kernel stores per node: it also keeps u in FP64 (a float
# potential stops following the resting state); nai, Ki, h and j are floats
# holding differences from their initial values
B_STORED = {"fp64": 304, "fp32": 180}
B_TRACK = 4
# the serving leg runs TP06 tissues at or above this fibrotic fraction on the
# FP64 path (activation times within one step of the FP64 path throughout)
FP64_FIBROSIS = 0.3


def profiled_traffic(key: str):
    """HBM bytes per launch of a kernel as measured with rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters (scaled by a device copy of known size)
    and recorded in profiles/*.md as `traffic_bytes_per_launch <key>: <B>` (or
    `traffic_bytes_per_step`, for a launch that advances several steps);
    (None, None) when no summary holds it."""
    for p in sorted((ROOT / "profiles").glob("*.md")):
        m = re.search(rf"^traffic_bytes_per_(?:launch|step) {key}: (\d+)\s*$", p.read_text(),
                      re.M)
        if m:
            return int(m.group(1)), f"profiles/{p.name}"
    return None, None


def build_model(precision: str):
    import finitewave_amd as fw
    n = N_SIDE + 2
    tissue = fw.CardiacTissue2D([n, n])
    ang = np.deg2rad(FIBER_DEG)
    fibers = np.zeros((n, n, 2))
    fibers[..., 0] = np.cos(ang)
    fibers[..., 1] = np.sin(ang)
    tissue.fibers = fibers
    model = fw.TP062D()
    model.dt = 0.01
    model.dr = DR
    model.t_max = 1e9
    model.prog_bar = False
    model.precision = precision
    model.cardiac_tissue = tissue
    seq = fw.StimSequence()
    seq.add_stim(fw.StimCurrentCoord2D(0, 100, 1.0, 0, n, 0, 11))
    model.stim_sequence = seq
    tracker = fw.ActivationTime2DTracker()
    tseq = fw.TrackerSequence()
    tseq.add_tracker(tracker)
    model.tracker_sequence = tseq
    return model, tracker


def _host_ahead(cycles: int = 4_000_000) -> None:
    """A wait kernel (~2 ms) on the current stream, enqueued after the
    synchronization that opens a timed window and before its start event:
    the host enqueues the window's launches while it runs, so the events
    time the launches' execution on the GPU, gaps between graphs included,
    as in a long run where the host keeps ahead of the device, and not the
    host's latency to its first launch."""
    import torch
    torch.cuda._sleep(cycles)


def gpu_run(precision: str, steps: int, warmup: int):
    """The S1 current's steps (0-100, on the device schedule, applied by the
    step kernels) untimed, then `warmup` untimed steps, then `steps` timed
    ones, as model.run() performs them, activation tracking on every step.
    The timed window starts after the stimulus has passed: C2's 100,000-step
    run spends all but 200 of its steps (S1 and S2) there, and the
    stimulated cells make a step ~11 % slower (DESIGN.md §6 gives both).
    The graphs the timed call replays are captured before the start event."""
    import torch
    from finitewave_amd.engine import Engine
    model, tracker = build_model(precision)
    model.initialize()
    eng = Engine(model)
    eng.upload()
    eng.act[tracker] = torch.from_numpy(tracker._steps_host()).to(eng.dev)
    stims = list(model.stim_sequence.sequence)
    model._schedule_stims(eng, stims, S1_END + 1 + warmup + steps)
    s1 = [(e.step_lo, e.step_hi) for e in eng.stim_sched]
    pre = max(hi for _, hi in s1) + 1  # the first step after the stimulus
    eng.run_steps(0, pre + warmup, tracker, np.ones(pre + warmup, np.uint8), use_graph=True)
    t0 = pre + warmup
    eng.prepare_graphs(steps, tracker)
    torch.cuda.synchronize()
    flags = np.ones(steps, np.uint8)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    _host_ahead()
    start.record()
    eng.run_steps(t0, steps, tracker, flags, use_graph=True)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = start.elapsed_time(end)
    u = eng.u_dev()
    finite = bool(torch.isfinite(u).all().item())
    act = eng.act[tracker]
    activated = int((act >= 0).sum().item())
    n_active = int(len(model.cardiac_tissue.myo_indexes))
    n_nodes = int(model.cardiac_tissue.mesh.size)
    stimulated = sum(max(0, min(hi, t0 + steps - 1) - max(lo, t0) + 1) for lo, hi in s1)
    eng.close()
    return dict(ms=ms, wall=wall, steps=steps, n_active=n_active, n_nodes=n_nodes,
                finite=finite, activated=activated, s1_steps=s1, stimulated_timed=stimulated,
                timed_steps=[t0, t0 + steps - 1])


def _engine_steps(model, tracker, steps: int, warmup: int):
    """Per-step time (us) of `steps` graph-replayed fused steps after
    `warmup` untimed ones, HIP events on the launch stream; activation
    tracking on every step."""
    import torch
    from finitewave_amd.engine import Engine
    model.initialize()
    eng = Engine(model)
    eng.upload()
    eng.act[tracker] = torch.from_numpy(tracker._steps_host()).to(eng.dev)
    eng.run_steps(0, warmup, tracker, np.ones(warmup, np.uint8), use_graph=True)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    _host_ahead()
    e0.record()
    eng.run_steps(warmup, steps, tracker, np.ones(steps, np.uint8), use_graph=True)
    e1.record()
    torch.cuda.synchronize()
    finite = bool(torch.isfinite(eng.u_dev()).all().item())
    eng.close()
    return e0.elapsed_time(e1) * 1e3 / steps, finite


def c1_leg():
    """The reference's CPU-runnable case: Aliev-Panfilov 2D, 100x100 tissue
    plus boundary, isotropic 5-point, FP64, planar stimulus on the 5 leftmost
    tissue columns, 1,000 steps (t_max 10), activation-time tracker.  Its
    state (0.4 MB) stays in L2: a latency figure, not a roofline one."""
    import torch
    import finitewave_amd as fw

    def model():
        m = fw.AlievPanfilov2D()
        m.dt, m.dr, m.t_max = 0.01, 0.25, 10
        m.prog_bar = False
        m.cardiac_tissue = fw.CardiacTissue2D([102, 102])
        seq = fw.StimSequence()
        seq.add_stim(fw.StimVoltageCoord2D(0, 1.0, 0, 102, 0, 6))
        m.stim_sequence = seq
        tr = fw.ActivationTime2DTracker()
        tr.threshold = 0.5
        ts = fw.TrackerSequence()
        ts.add_tracker(tr)
        m.tracker_sequence = ts
        return m, tr

    m, tr = model()
    us, finite = _engine_steps(m, tr, 1000, 100)
    m, tr = model()
    m.run()
    torch.cuda.synchronize()
    walls = []
    for _ in range(3):
        w0 = time.perf_counter()
        m.run()
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - w0)
    return dict(workload="Aliev-Panfilov 2D 102x102 (100x100 tissue), isotropic 5-point, FP64, "
                         "planar stimulus, activation-time tracker",
                us_per_step=round(us, 3), gcell=round(10_000 / us / 1e3, 3),
                run_1000_steps_ms=round(min(walls) * 1e3, 2), finite=finite)


def _tip_track(df):
    """Per tracked time, the centroid of the spiral tips found then, in
    nodes: {time: (x, y)}."""
    out = {}
    if df is None or len(df) == 0:
        return out
    for t, g in df.groupby("time"):
        out[round(float(t), 6)] = (float(g["x"].mean()), float(g["y"].mean()))
    return out


def _c2_spiral_run(prec: str, s2_time: float, t_max: float, cells):
    import torch
    import finitewave_amd as fw
    n = N_SIDE + 2
    model, act = build_model(prec)
    model.t_max = t_max
    model.stim_sequence.add_stim(fw.StimCurrentCoord2D(s2_time, 100, 1.0, n // 2, n, 0, n // 2))
    act.threshold = -40
    tips = fw.SpiralWaveCore2DTracker()
    tips.threshold = -40
    tips.start_time = s2_time + 10
    tips.step = 100
    per = fw.Period2DTracker()
    per.cell_ind = cells
    per.threshold = -40
    per.step = 10
    model.tracker_sequence.add_tracker(tips)
    model.tracker_sequence.add_tracker(per)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    model.run()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    periods = [np.asarray(p, dtype=np.float64) for p in per.output]
    out = tips.output
    print(f"# c2 spiral S2 {s2_time} {prec}: {model.step} steps in {wall:.1f} s, "
          f"{len(out)} tips", file=sys.stderr, flush=True)
    return dict(act=np.asarray(act.output, dtype=np.float64), tips=_tip_track(out),
                n_tips=int(len(out)), periods=periods, wall=wall, steps=int(model.step),
                finite=bool(np.all(np.isfinite(model.u))))


def c2_spiral_leg(s2_times=(310.0, 340.0), t_max: float = 1000.0):
    """C2 at its full length: TP06 1026^2, 30 deg fibres, S1 planar on the
    left 10 columns at t = 0, S2 on the lower-left quadrant, 100,000 steps,
    run by model.run() in FP64 and in FP32 (FP64 calcium) with three
    trackers: activation time (threshold -40 mV), spiral tips
    (SpiralWaveCore2DTracker, every 1 ms from S2 + 10 ms) and activation
    periods at five detectors (Period2DTracker, every 0.1 ms).  S2 at 310 ms
    as configured, and at 340 ms: at 310 ms the S1 wave has not yet crossed
    the S2 quadrant's edge and the S2 wave dies out (no tip is ever found);
    340 ms is the earliest S2 of 310, 340, 370, 400 and 430 ms that leaves a
    rotating wave.  Spiral dynamics diverge chaotically, so the precisions
    are compared by statistics: the S1 front's activation times (nodes
    activated before S2) within one dt, the tip trajectories (distance of
    the FP64 and FP32 tip centroids over time, and the time they first part
    by 5 nodes) and the median activation period at the detectors."""
    n = N_SIDE + 2
    cells = [[n // 4, n // 4], [n // 2, n // 2], [3 * n // 4, n // 4], [n // 4, 3 * n // 4],
             [3 * n // 4, 3 * n // 4]]
    dt = 0.01
    legs = {}
    for s2 in s2_times:
        runs = {prec: _c2_spiral_run(prec, s2, t_max, cells) for prec in ("fp64", "fp32")}
        a, b = runs["fp64"]["act"], runs["fp32"]["act"]
        s1 = (a >= 0) & (a < s2)
        both = s1 & (b >= 0)
        dlat = np.abs(a[both] - b[both]) if both.any() else np.zeros(1)
        t64, t32 = runs["fp64"]["tips"], runs["fp32"]["tips"]
        common = sorted(set(t64) & set(t32))
        dist = np.array([math.hypot(t64[t][0] - t32[t][0], t64[t][1] - t32[t][1])
                         for t in common])
        apart = [t for t, d in zip(common, dist) if d > 5.0]

        def med_period(r):
            vals = [float(np.median(p[p > 0])) for p in r["periods"] if np.any(p > 0)]
            return round(float(np.median(vals)), 2) if vals else None

        legs[f"s2_{int(s2)}ms"] = dict(
            steps={k: r["steps"] for k, r in runs.items()},
            wall_s={k: round(r["wall"], 2) for k, r in runs.items()},
            ms_per_step={k: round(r["wall"] / max(1, r["steps"]) * 1e3, 4)
                         for k, r in runs.items()},
            finite={k: r["finite"] for k, r in runs.items()},
            s1_front=dict(nodes=int(s1.sum()), both_activated=int(both.sum()),
                          missing_in_fp32=int((s1 & (b < 0)).sum()),
                          lat_max_dt=round(float(dlat.max()) / dt, 3),
                          beyond_1dt=int(np.count_nonzero(dlat > dt + 1e-9))),
            tips=dict(found={k: r["n_tips"] for k, r in runs.items()},
                      tracked_times_with_tips={k: len(r["tips"]) for k, r in runs.items()},
                      common_times=len(common),
                      centroid_dist_nodes_p50_p90=[round(float(x), 2) for x in
                                                   np.percentile(dist, [50, 90])]
                      if len(dist) else None,
                      first_apart_5_nodes_ms=apart[0] if apart else None,
                      last_common_ms=common[-1] if common else None),
            median_period_ms={k: med_period(r) for k, r in runs.items()})
    return dict(workload="TP06 2D 1024x1024 (+boundary), 9-point, fibre 30 deg, S1 planar t=0, "
                         "S2 lower-left quadrant, 100,000 steps, model.run() with activation, "
                         "spiral-tip and period trackers, FP64 and FP32",
                **legs)


def c3_leg(steps: int = 400, warmup: int = 40):
    """Aliev-Panfilov 3D slab, 256^3 tissue plus boundary, anisotropic
    19-point stencil, fibres in the xy-plane rotating from -60 deg (first
    tissue layer in z) to +60 deg (last), D_long 1.0, D_trans 0.111, dr 0.25,
    dt 0.01, diffuse fibrosis i.i.d. p = 0.2 (numpy seed 0, the Diffuse3D
    pattern), FP32, planar stimulus on the x < 5 face, activation-time
    tracker.  Roofline: (u, v read and written in float + the 4-byte
    activation word) per tissue node + the 1-byte code per node."""
    import finitewave_amd as fw
    n = 258
    tissue = fw.CardiacTissue3D([n, n, n])
    np.random.seed(0)
    tissue.add_pattern(fw.Diffuse3DPattern(0, n, 0, n, 0, n, 0.2))
    k = np.arange(n)
    phi = np.deg2rad(-60 + 120 * np.clip(k - 1, 0, n - 3) / (n - 3))
    f = np.zeros((n, n, n, 3))
    f[..., 0] = np.cos(phi)[None, None, :]
    f[..., 1] = np.sin(phi)[None, None, :]
    tissue.fibers = f
    m = fw.AlievPanfilov3D()
    m.dt, m.dr, m.t_max = 0.01, 0.25, 1e9
    m.prog_bar = False
    m.precision = "fp32"
    m.cardiac_tissue = tissue
    st = fw.AsymmetricStencil3D()
    st.D_al, st.D_ac = 1.0, 0.111
    m.stencil = st
    seq = fw.StimSequence()
    seq.add_stim(fw.StimVoltageCoord3D(0, 1.0, 0, 6, 0, n, 0, n))
    m.stim_sequence = seq
    tr = fw.ActivationTime3DTracker()
    tr.threshold = 0.5
    ts = fw.TrackerSequence()
    ts.add_tracker(tr)
    m.tracker_sequence = ts
    us, finite = _engine_steps(m, tr, steps, warmup)
    na = int(np.count_nonzero(tissue.mesh == 1))
    nodes = int(tissue.mesh.size)
    bytes_step = (16 + B_TRACK) * na + nodes
    achieved = bytes_step / (us * 1e-6) / 1e9
    traffic, source = profiled_traffic("c3")
    return dict(workload="Aliev-Panfilov 3D 258^3 (256^3 tissue), anisotropic 19-point, fibres "
                         "-60..60 deg along z, fibrosis p=0.2, FP32, activation-time tracker",
                nodes=nodes, active=na, steps_timed=steps, us_per_step=round(us, 2),
                gcell=round(na / us / 1e3, 3), finite=finite,
                roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                          "traffic": traffic, "traffic_source": source,
                          "bytes_per_step": bytes_step})


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(steps: int, threads: int):
    """The oracle (C restatement of the reference's numba kernels, same data
    layout: (*shape, K) FP64 weights, int64 active-index list, one diffusion
    and one ionic sweep per step) on `threads` host threads: the headline
    configuration (a bounded sample of TP06 1026^2 steps), the reference's own
    CPU case in full (Aliev-Panfilov 102^2, 1,000 steps), and samples of the
    3D slab and of the restitution ensemble from which their rates follow."""
    from oracle.oracle import OracleSim, OStim, set_threads
    used = set_threads(threads)
    n = N_SIDE + 2
    mesh = np.ones((n, n), dtype=np.int8)
    mesh[0, :] = mesh[-1, :] = mesh[:, 0] = mesh[:, -1] = 0
    ang = np.deg2rad(FIBER_DEG)
    fibers = np.zeros((n, n, 2))
    fibers[..., 0] = np.cos(ang)
    fibers[..., 1] = np.sin(ang)
    sim = OracleSim("tp06", mesh, 0.01, DR, fibers=fibers)
    sim.stims.append(OStim("current", 0, 100, (slice(0, n), slice(0, 11)), duration=1.0))
    sim.initialize()
    sim.one_step()  # warm caches and the thread pool
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.one_step()
        sim.u, sim.u_new = sim.u_new, sim.u
    dt = time.perf_counter() - t0
    cells = len(sim.idx) * steps
    legs = {}
    # C1 in full: the oracle's run loop (stimulus, sweeps, activation tracker)
    from oracle.oracle import OTracker
    m1 = np.ones((102, 102), dtype=np.int8)
    m1[0, :] = m1[-1, :] = m1[:, 0] = m1[:, -1] = 0
    c1 = OracleSim("aliev_panfilov", m1, 0.01, 0.25)
    c1.stims.append(OStim("voltage", 0, 1.0, (slice(0, 102), slice(0, 6))))
    c1.trackers.append(OTracker("act_time", threshold=0.5))
    w0 = time.perf_counter()
    c1.run(10.0)
    w1 = time.perf_counter() - w0
    legs["c1"] = dict(steps=c1.step, run_ms=round(w1 * 1e3, 2),
                      us_per_step=round(w1 / c1.step * 1e6, 2),
                      gcell=round(10_000 * c1.step / w1 / 1e9, 4))
    # C3 sample: the 258^3 slab in FP64 (the reference's arithmetic)
    k3 = 258
    rs = np.random.RandomState(0)
    m3 = (1 + (rs.random_sample((k3,) * 3) <= 0.2)).astype(np.int8)
    m3[0], m3[-1], m3[:, 0], m3[:, -1], m3[:, :, 0], m3[:, :, -1] = 0, 0, 0, 0, 0, 0
    kk = np.arange(k3)
    phi = np.deg2rad(-60 + 120 * np.clip(kk - 1, 0, k3 - 3) / (k3 - 3))
    f3 = np.zeros((k3, k3, k3, 3))
    f3[..., 0] = np.cos(phi)[None, None, :]
    f3[..., 1] = np.sin(phi)[None, None, :]
    c3 = OracleSim("aliev_panfilov", m3, 0.01, 0.25, fibers=f3, stencil="aniso", D_ac=0.111)
    c3.initialize()
    del f3
    c3.one_step()
    n3 = 3
    w0 = time.perf_counter()
    for _ in range(n3):
        c3.one_step()
        c3.u, c3.u_new = c3.u_new, c3.u
    w3 = time.perf_counter() - w0
    legs["c3"] = dict(steps=n3, gcell=round(len(c3.idx) * n3 / w3 / 1e9, 4),
                      ms_per_step=round(w3 / n3 * 1e3, 1))
    del c3
    # C4 sample: one 256^2 Mitchell-Schaeffer tissue of the ensemble; the
    # ensemble's rate follows from its total cell-steps
    from finitewave_amd.workloads import restitution_ensemble
    _, sims = restitution_ensemble(n=512)
    total = sum(int(np.count_nonzero(s.tissue.mesh == 1)) * int(math.ceil(s.t_max / 0.01))
                for s in sims)
    pick = [s for s in sims if s.tag["side"] == 256][0]
    c4 = OracleSim("mitchell_schaeffer", pick.tissue.mesh, 0.01, 0.25)
    c4.initialize()
    c4.one_step()
    n4 = 400
    w0 = time.perf_counter()
    for _ in range(n4):
        c4.one_step()
        c4.u, c4.u_new = c4.u_new, c4.u
    w4 = time.perf_counter() - w0
    rate4 = len(c4.idx) * n4 / w4
    legs["c4"] = dict(sample="400 steps of one 256^2 tissue", gcell=round(rate4 / 1e9, 4),
                      projected_wall_s=round(total / rate4, 1),
                      sims_per_s=round(512 / (total / rate4), 5))
    return dict(value=cells / dt / 1e9, unit="Gcell-steps/s", cores=used, kind="port",
                cpu=_cpu_model(),
                sample=f"{steps} FP64 steps of the same TP06 1026x1026 aniso workload, "
                       f"{dt:.1f} s on {used} threads",
                legs=legs)


def ensemble_leg(n_tissues: int, steps: int, full: bool):
    """The restitution ensemble (Mitchell-Schaeffer, FP32): every tissue is
    admitted into one batch, the S1 beat is applied, and `steps` steps of
    the whole batch are timed (all tissues are present for the whole S1
    train, so the per-step cost is stationary).  With `full` the ensemble is
    also run to completion and its sims/s measured."""
    import torch
    from finitewave_amd.batch import TissueBatch
    from finitewave_amd.workloads import restitution_ensemble
    cls, sims = restitution_ensemble(n=n_tissues)
    b = TissueBatch(cls, precision="fp32", max_tissues=n_tissues, chunk_cap=4000)
    t_adm = time.perf_counter()
    for s in sims:
        b.admit(s)
    torch.cuda.synchronize()
    t_adm = time.perf_counter() - t_adm
    # S1 at step 0; the autotuner times every launch shape on the batch's
    # first steps; then the graphs of the chosen shape are captured: all
    # before the timed window
    b.advance((b.tune_steps + 1) * len(b.variants()) + 200)
    torch.cuda.synchronize()
    # events on the batch's own stream, the one its kernels are launched on
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(b.stream)
    done = 0
    while done < steps:
        n0 = b.step
        b.advance(steps - done)
        done += b.step - n0
    ev1.record(b.stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    n_active = sum(int(np.count_nonzero(s.tissue.mesh == 1)) for s in sims)
    n_nodes = sum(int(s.tissue.mesh.size) for s in sims)
    sec = ms / 1e3 / done
    # MS, FP32: u, u_new and h (read + write) per active node, the node code per node
    bytes_step = 16 * n_active + n_nodes
    total_cell_steps = sum(int(np.count_nonzero(s.tissue.mesh == 1)) *
                           int(math.ceil(s.t_max / 0.01)) for s in sims)
    out = dict(workload=f"{n_tissues} Mitchell-Schaeffer 2D tissues, side U{{64,128,256,512}}+2, "
                        "fibrosis p~U[0,0.35], S1 x5 @ 400 ms + S2 @ CI~U[150,400] ms, "
                        "isotropic, dt 0.01, dr 0.25, FP32",
               tissues=n_tissues, nodes=n_nodes, active=n_active, steps_timed=done,
               ms_per_step=round(sec * 1e3, 5), gcell=round(n_active / sec / 1e9, 3),
               admit_s=round(t_adm, 3),
               roofline={"bound": "hbm", "achieved": round(bytes_step / sec / 1e9, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(bytes_step / sec / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": profiled_traffic("ensemble")[0] if n_tissues == 512 else None,
                         "traffic_source": profiled_traffic("ensemble")[1],
                         "bytes_per_step": bytes_step},
               projected_s=round(total_cell_steps / (n_active / sec), 1),
               launch_variant=b.variant,
               steps_per_launch=(2 << ((b.variant >> 4) & 3)) if b.variant & 8 else 1,
               tuning_ms={str(k): round(v, 4) for k, v in (b.tuning or {}).items()})
    if full:
        w0 = time.perf_counter()
        b.run_all()
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0 + done * sec
        res = [s.result for s in sims]
        apds = [r["apd90"] for r in res if r.get("apd90") is not None]
        cvs = [r["cv_front"] for r in res if r.get("cv_front") is not None]
        out["full"] = dict(wall_s=round(wall, 2), sims_per_s=round(n_tissues / wall, 3),
                           captured=sum(bool(r.get("captured")) for r in res),
                           loss_of_capture=sum(not r.get("captured") for r in res),
                           apd90_found=len(apds),
                           apd90_ms_p10_p50_p90=[round(float(x), 2) for x in
                                                 np.percentile(apds, [10, 50, 90])] if apds else None,
                           cv_front_found=len(cvs),
                           cv_front_p10_p50_p90=[round(float(x), 4) for x in
                                                 np.percentile(cvs, [10, 50, 90])] if cvs else None,
                           cv_row_found=sum(r.get("cv") is not None for r in res))
    b.close()
    return out


def _fp64_plain(cls, sim):
    """One request on the plain FP64 single-tissue path (model.run())."""
    import copy
    import finitewave_amd as fw
    model = cls()
    model.precision = "fp64"
    model.dt, model.dr, model.t_max = 0.01, 0.25, sim.t_max
    model.prog_bar = False
    model.cardiac_tissue = copy.deepcopy(sim.tissue)
    seq = fw.StimSequence()
    for s in sim.stims:
        seq.add_stim(copy.deepcopy(s))
    model.stim_sequence = seq
    ts = fw.TrackerSequence()
    act = fw.ActivationTime2DTracker()
    act.threshold = sim.act_threshold
    ts.add_tracker(act)
    ap = fw.ActionPotential2DTracker()
    ap.cell_ind = list(sim.probe)
    ap.step = 1
    ts.add_tracker(ap)
    model.tracker_sequence = ts
    model.run()
    return np.array(act.output), np.array(ap.output)


def serving_leg(n_req: int, fp64_baseline: bool, per_request: bool = False,
                plain_every: int = 4):
    """The Poisson request stream served with continuous batching (FP32,
    TP06 calcium in FP64), then, for comparison and for the error figures,
    every `plain_every`-th request again, one after another, on the plain
    FP64 path (its sims/s is that sample's rate)."""
    import torch
    from finitewave_amd.analysis import apd90
    from finitewave_amd.serving import Request, Server, fp64_for_fibrosis
    from finitewave_amd.workloads import serving_stream
    classes, specs = serving_stream(n=n_req)
    reqs = [Request(rid=k, arrival=s["arrival"], model=s["model"], sim=s["sim"])
            for k, s in enumerate(specs)]
    # FP32, except TP06 tissues close to conduction failure: there float
    # rounding moves activation times by more than a step
    policy = fp64_for_fibrosis(("tp06",), FP64_FIBROSIS)
    srv = Server(classes, precision="fp32", max_tissues=64, chunk_steps=1000,
                 precision_for=policy)

    def served(r):  # progress: one line per finished request
        print(f"# serving: request {r.rid} ({r.model}, {r.precision}) {r.status} at "
              f"{r.finished:.1f} s", file=sys.stderr, flush=True)
    stats = srv.serve(reqs, realtime=True, log=served)
    shapes = {(m if p == "fp32" else f"{m}_{p}"): dict(
                  variant=b.variant,
                  tuning_ms={str(v): round(t, 4) for v, t in (b.tuning or {}).items()})
              for (m, p), b in srv.batches.items()}
    srv.close()
    cell_steps = sum(int(np.count_nonzero(r.sim.tissue.mesh == 1)) * r.sim.n_steps for r in reqs)
    out = dict(workload=f"{n_req} requests, Poisson 200/s, TP06 or Aliev-Panfilov 2D, side "
                        "U{128..1024}+2, fibrosis p~U[0,0.4], fibre U[0,180) deg (9-point), "
                        "planar or S1-S2 cross-field, horizon U[500,2000] ms, FP32 (TP06 Ca "
                        "FP64); TP06 tissues with fibrotic fraction >= "
                        f"{FP64_FIBROSIS} on the FP64 path",
               requests=n_req, sims_per_s=round(stats["sims_per_s"], 4),
               fp64_requests=sum(r.precision == "fp64" for r in reqs),
               p50_s=round(stats["p50_s"], 3), p99_s=round(stats["p99_s"], 3),
               wall_s=round(stats["wall_s"], 2),
               gcell=round(cell_steps / stats["wall_s"] / 1e9, 3),
               max_concurrent=stats["max_concurrent"],
               peak_hbm_mb=round(stats["peak_hbm_bytes"] / 2**20, 1),
               launch_shapes=shapes)
    if fp64_baseline:
        torch.cuda.synchronize()
        t_plain = 0.0
        lat_max, lat_bad, lat_n, apd_err = 0.0, 0, 0, []
        per_req = []
        sample = reqs[::plain_every]
        for r in sample:
            w0 = time.perf_counter()
            act64, trace64 = _fp64_plain(classes[r.model], r.sim)
            torch.cuda.synchronize()
            t_plain += time.perf_counter() - w0
            print(f"# fp64 plain: request {r.rid} ({r.model}) in "
                  f"{time.perf_counter() - w0:.1f} s", file=sys.stderr, flush=True)
            a32 = r.sim.result["act_t"]
            both = (a32 >= 0) & (act64 >= 0)
            d = np.abs(a32[both] - act64[both])
            lat_n += int(both.sum())
            lat_bad += int(np.count_nonzero(d > 0.01 + 1e-9)) + \
                int(np.count_nonzero((a32 >= 0) != (act64 >= 0)))
            lat_max = max(lat_max, float(d.max()) if d.size else 0.0)
            a = r.sim.result.get("apd90")
            a64 = apd90(trace64, 0.01, r.sim.apd_threshold)
            if a is not None and a64 is not None:
                apd_err.append(abs(a - a64))
            tag = r.sim.tag
            per_req.append(dict(model=r.model, precision=r.precision, side=tag["side"],
                                p=round(tag["p"], 3),
                                cross=tag["cross"], lat_max_ms=float(d.max()) if d.size else 0.0,
                                lat_p99_ms=float(np.percentile(d, 99)) if d.size else 0.0,
                                over_1dt=int(np.count_nonzero(d > 0.01 + 1e-9)),
                                only_one=int(np.count_nonzero((a32 >= 0) != (act64 >= 0))),
                                nodes=int(both.sum()),
                                apd90_err_ms=(abs(a - a64) if a is not None and a64 is not None
                                              else None)))
        plain_rate = len(sample) / t_plain
        # the served rate over the same requests, for a like-for-like ratio
        out["fp64_plain"] = dict(sims_per_s=round(plain_rate, 4), wall_s=round(t_plain, 2),
                                 requests=len(sample),
                                 sample=f"every {plain_every}th request of the stream",
                                 cell_steps_per_s_g=round(sum(
                                     int(np.count_nonzero(r.sim.tissue.mesh == 1)) * r.sim.n_steps
                                     for r in sample) / t_plain / 1e9, 3))
        # the plain path's time for the whole stream, projected from the
        # sample by cell-steps (its cost is the step kernels')
        cs_sample = sum(int(np.count_nonzero(r.sim.tissue.mesh == 1)) * r.sim.n_steps
                        for r in sample)
        plain_all = t_plain * cell_steps / cs_sample
        out["fp64_plain"]["projected_stream_wall_s"] = round(plain_all, 1)
        out["fp64_plain"]["projected_sims_per_s"] = round(n_req / plain_all, 4)
        out["speedup_vs_fp64_plain"] = round(stats["sims_per_s"] / (n_req / plain_all), 2)
        out["errors_vs_fp64"] = dict(lat_max_ms=lat_max, lat_nodes=lat_n,
                                     lat_over_1dt=lat_bad,
                                     lat_frac_over_1dt=lat_bad / max(1, lat_n),
                                     lat_p99_ms_worst_request=max(
                                         (q["lat_p99_ms"] for q in per_req), default=None),
                                     apd90_max_ms=max(apd_err) if apd_err else None,
                                     apd90_compared=len(apd_err))
        if per_request:
            out["per_request"] = per_req
    return out


def copy_calibration(mbytes: int = 1024) -> None:
    """A plain device copy of known size, for scaling the HBM counters."""
    import torch
    n = mbytes * (1 << 20) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32"], default=None)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--ensemble-tissues", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=2000)
    ap.add_argument("--no-ensemble-full", action="store_true",
                    help="skip running the whole ensemble to completion")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--plain-every", type=int, default=4,
                    help="run every k-th served request on the plain FP64 path")
    ap.add_argument("--no-fp64-serving", action="store_true")
    ap.add_argument("--per-request", action="store_true",
                    help="list each served request's errors against the FP64 path")
    ap.add_argument("--legs", default="stencil,ensemble,serving,c1,c2spiral,c3")
    args = ap.parse_args(argv)
    legs = set(args.legs.split(","))
    if args.gpus != 1:
        print("bench.py measures one GPU: the stepping path has no multi-GPU mode (the slab "
              "decomposition across GPUs is designed, not built); run with --gpus 1",
              file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration()
    runs = {}
    for prec in (["fp64", "fp32"] if args.only is None else [args.only]):
        runs[prec] = gpu_run(prec, args.steps, args.warmup)
        r = runs[prec]
        print(f"# {prec}: {r['ms'] / r['steps'] * 1e3:.2f} us/step, finite={r['finite']}, "
              f"activated={r['activated']}, S1 on steps {r['s1_steps']}, "
              f"{r['stimulated_timed']} of the timed steps stimulated", file=sys.stderr,
              flush=True)
    primary = "fp64" if "fp64" in runs else "fp32"
    out = {}
    for prec, r in runs.items():
        sec_per_step = r["ms"] / 1e3 / r["steps"]
        bytes_step = (B_STATE[prec] + B_TRACK) * r["n_active"] + r["n_nodes"]
        achieved = bytes_step / sec_per_step / 1e9
        stored = (B_STORED[prec] + B_TRACK) * r["n_active"] + r["n_nodes"]
        out[prec] = dict(
            gcell=r["n_active"] / sec_per_step / 1e9,
            ms_per_step=sec_per_step * 1e3,
            roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                      "traffic": profiled_traffic(prec)[0],
                      "traffic_source": profiled_traffic(prec)[1],
                      "bytes_per_step": bytes_step,
                      "stored_bytes_per_step": stored,
                      "stored_frac": round(stored / sec_per_step / 1e9 / HBM_PEAK_GBS, 4)},
            finite=r["finite"], activated=r["activated"],
            stimulated_timed=r["stimulated_timed"], s1_steps=r["s1_steps"])
    p = out[primary]
    line = {
        "metric": "Gcell-steps/s & %HBM roofline; served sims/s + p50/p99 at LAT err <= 1 dt",
        "value": round(p["gcell"], 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(p["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if primary == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "TP06 2D 1024x1024 (+boundary), anisotropic 9-point, fibre 30 deg, "
                               "dr 0.025 cm (0.25 model units), dt 0.01, S1 planar stimulus, "
                               "activation-time tracker",
                   "nodes": runs[primary]["n_nodes"], "active": runs[primary]["n_active"],
                   "precision": primary,
                   "timed_steps": runs[primary]["timed_steps"],
                   "untimed_before": "the S1 stimulus steps, then the warm-up steps",
                   "s1_stimulus_steps": list(runs[primary]["s1_steps"][0]),
                   "stimulated_timed_steps": runs[primary]["stimulated_timed"]},
        "roofline": p["roofline"],
    }
    if "fp32" in out and primary != "fp32":
        f = out["fp32"]
        line["fp32_mixed"] = {"value": round(f["gcell"], 4), "unit": "Gcell-steps/s",
                              "ms_per_step": round(f["ms_per_step"], 5),
                              "roofline": f["roofline"],
                              "stimulated_timed_steps": f["stimulated_timed"]}
    if "ensemble" in legs:
        line["ensemble"] = ensemble_leg(args.ensemble_tissues, args.ensemble_steps,
                                        not args.no_ensemble_full)
        print(f"# ensemble: {line['ensemble']}", file=sys.stderr, flush=True)
    if "serving" in legs:
        line["serving"] = serving_leg(args.serve_requests, not args.no_fp64_serving,
                                      args.per_request, args.plain_every)
        print(f"# serving: {line['serving']}", file=sys.stderr, flush=True)
    if "c1" in legs:
        line["c1"] = c1_leg()
        print(f"# c1: {line['c1']}", file=sys.stderr, flush=True)
    if "c2spiral" in legs:
        line["c2_spiral"] = c2_spiral_leg()
        print(f"# c2 spiral: {line['c2_spiral']}", file=sys.stderr, flush=True)
    if "c3" in legs:
        line["c3"] = c3_leg()
        print(f"# c3: {line['c3']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(args.cpu_steps, threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
