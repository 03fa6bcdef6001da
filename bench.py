"""Benchmark of the fused stepping kernel on one MI355X.

Workload (the reference's headline configuration that fits one GPU): TP06
epicardial, 1024x1024 tissue plus a one-node zero boundary (1026x1026),
anisotropic 9-point stencil with a uniform fibre angle of 30 degrees
(D_long 0.154, D_trans 0.154/9), dr 0.025 cm, dt 0.01 ms, planar S1
current stimulus on the left 10 tissue columns at t = 0, activation-time
tracker on.  A step is one fused diffusion + ionic + tracker pass over the
whole tissue; inputs are resident in HBM when timing starts.

Units: the reference's TP06 takes D_model = 0.154 in mm^2/ms and dr in mm
(its own example, examples/models/2D/tp06_2d.py:47, sets dr = 0.25), so
dr = 0.025 cm enters the model as dr = 0.25.  Entered as 0.025 the explicit
scheme has D*dt/dr^2 = 2.46 > 1/4 and diverges, in the reference as here.

The JSON line's value is the FP64 path (the reference's arithmetic); the
FP32 path (TP06 calcium and concentrations in FP64) is reported beside it.
roofline: algorithmic bytes per step / measured time per step against the
8 TB/s HBM3E peak.  cpu_baseline: the C/OpenMP oracle on the host cores over
a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
N_SIDE = 1024
FIBER_DEG = 30.0
DR = 0.25  # 0.025 cm in the model's millimetres

# bytes each step must move per active node: state read + state written
# (TP06 reads u and 18 state arrays and writes u_new and 18 arrays; FP32 path:
# 14 gates/u as FP32, cai/casr/cass/nai/Ki as FP64), plus the 1-byte node
# code per node and the 4-byte activation-time word per node
B_STATE = {"fp64": 304, "fp32": 188}
B_TRACK = 4


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


def gpu_run(precision: str, steps: int, warmup: int):
    import torch
    from finitewave_amd.engine import Engine
    model, tracker = build_model(precision)
    model.initialize()
    eng = Engine(model)
    eng.upload()
    eng.act[tracker] = torch.from_numpy(tracker._steps_host()).to(eng.dev)
    stim = model.stim_sequence.sequence[0]
    # warm-up: the S1 stimulus window then plain steps, all untimed
    t = 0.0
    s = 0
    while s < warmup:
        if t >= stim.t and not stim.passed:
            stim._apply_device(model, eng)
            model.t = t
            stim.update_status(model)
            eng.run_steps(s, 1, tracker, np.ones(1, np.uint8), use_graph=False)
            s += 1
            t += model.dt
            continue
        n = warmup - s
        eng.run_steps(s, n, tracker, np.ones(n, np.uint8), use_graph=True)
        for _ in range(n):
            t += model.dt
        s += n
    torch.cuda.synchronize()
    flags = np.ones(steps, np.uint8)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    start.record()
    eng.run_steps(s, steps, tracker, flags, use_graph=True)
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
    eng.close()
    return dict(ms=ms, wall=wall, steps=steps, n_active=n_active, n_nodes=n_nodes,
                finite=finite, activated=activated)


def cpu_baseline(steps: int, threads: int):
    """The oracle (C restatement of the reference's numba kernels, same data
    layout: (*shape, 9) FP64 weights, int64 active-index list, one diffusion
    and one ionic sweep per step) on `threads` host threads."""
    from oracle.oracle import OracleSim, OStim, set_threads
    n = N_SIDE + 2
    mesh = np.ones((n, n), dtype=np.int8)
    mesh[0, :] = mesh[-1, :] = mesh[:, 0] = mesh[:, -1] = 0
    ang = np.deg2rad(FIBER_DEG)
    fibers = np.zeros((n, n, 2))
    fibers[..., 0] = np.cos(ang)
    fibers[..., 1] = np.sin(ang)
    used = set_threads(threads)
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
    return dict(value=cells / dt / 1e9, unit="Gcell-steps/s", cores=used, kind="port",
                sample=f"{steps} FP64 steps of the same TP06 1026x1026 aniso workload, "
                       f"{dt:.1f} s on {used} threads")


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
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: every MI355X box here has one device and the "
              "stepping path has no multi-GPU mode (slab decomposition is a design note); "
              "run with --gpus 1", file=sys.stderr)
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
              f"activated={r['activated']}", file=sys.stderr, flush=True)
    primary = "fp64" if "fp64" in runs else "fp32"
    out = {}
    for prec, r in runs.items():
        sec_per_step = r["ms"] / 1e3 / r["steps"]
        bytes_step = (B_STATE[prec] + B_TRACK) * r["n_active"] + r["n_nodes"]
        achieved = bytes_step / sec_per_step / 1e9
        out[prec] = dict(
            gcell=r["n_active"] / sec_per_step / 1e9,
            ms_per_step=sec_per_step * 1e3,
            roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                      "traffic": None, "bytes_per_step": bytes_step},
            finite=r["finite"], activated=r["activated"])
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
                   "precision": primary},
        "roofline": p["roofline"],
    }
    if "fp32" in out and primary != "fp32":
        f = out["fp32"]
        line["fp32_mixed"] = {"value": round(f["gcell"], 4), "unit": "Gcell-steps/s",
                              "ms_per_step": round(f["ms_per_step"], 5),
                              "roofline": f["roofline"]}
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(args.cpu_steps, threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
