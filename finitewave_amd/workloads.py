"""Synthetic workloads of the ensemble and serving benchmarks.

* ``restitution_ensemble``: Mitchell-Schaeffer 2D tissues of side drawn from
  {64, 128, 256, 512} (+2 boundary), i.i.d. fibrosis p ~ U[0, 0.35] drawn as
  1 + (random <= p) with RandomState(i) (the Diffuse pattern's form), an S1
  train of 5 planar voltage beats every 400 ms on the first 5 tissue columns,
  an S2 at 1600 ms + CI, CI ~ U[150, 400] ms, horizon S2 + 500 ms;
  activation recorded from S2 on, APD90 at the tissue centre on the trace
  from S2 on (no upstroke there = loss of capture), conduction velocity as
  the reference's front velocity and along the centre row.
* ``serving_stream``: Poisson arrivals (rate 200/s), TP06 or Aliev-Panfilov
  with probability 1/2, side ~ U{128..1024}, fibrosis p ~ U[0, 0.4], uniform
  fibre angle ~ U[0, 180) (9-point stencil), planar or S1-S2 cross-field
  stimulation, horizon ~ U[500, 2000] ms; results = activation map + APD90.
"""
from __future__ import annotations

import numpy as np

from . import stimulation as S
from .batch import BatchSim
from .models import AlievPanfilov2D, MitchellSchaeffer2D, TP062D
from .tissue import CardiacTissue2D


def fibrotic_tissue(side: int, p: float, seed: int, fibre_deg=None) -> CardiacTissue2D:
    n = side + 2
    t = CardiacTissue2D([n, n])
    rng = np.random.RandomState(seed)
    mesh = (1 + (rng.random_sample((n, n)) <= p)).astype(np.int8) if p > 0 else \
        np.ones((n, n), dtype=np.int8)
    t.mesh = mesh
    if fibre_deg is not None:
        a = np.deg2rad(fibre_deg)
        f = np.zeros((n, n, 2))
        f[..., 0], f[..., 1] = np.cos(a), np.sin(a)
        t.fibers = f
    return t


def _tissue_cell(mesh, i, j):
    """(i, j) or the nearest tissue node on row i."""
    row = np.flatnonzero(mesh[i] == 1)
    if len(row) == 0:
        return (i, j)
    return (i, int(row[np.argmin(np.abs(row - j))]))


def restitution_ensemble(n: int = 512, seed: int = 1234, sides=(64, 128, 256, 512),
                         beats: int = 5, bcl: float = 400.0, tail: float = 500.0,
                         p_range=(0.0, 0.35), ci_range=(150.0, 400.0)):
    rng = np.random.default_rng(seed)
    side = rng.choice(np.asarray(sides), size=n)
    p = rng.uniform(*p_range, size=n)
    ci = rng.uniform(*ci_range, size=n)
    sims = []
    for i in range(n):
        s = int(side[i])
        tissue = fibrotic_tissue(s, float(p[i]), i)
        s2 = (beats - 1) * bcl + float(ci[i])
        stims = [S.StimVoltageCoord2D(k * bcl, 1.0, 0, s + 2, 0, 6) for k in range(beats)]
        stims.append(S.StimVoltageCoord2D(s2, 1.0, 0, s + 2, 0, 6))
        mid = (s + 2) // 2
        sims.append(BatchSim(tissue=tissue, t_max=s2 + tail, stims=stims,
                             probe=_tissue_cell(tissue.mesh, mid, mid), act_threshold=0.5,
                             act_start=s2, apd_threshold=0.5, apd_start=s2, cv_row=mid,
                             tag=dict(side=s, p=float(p[i]), ci=float(ci[i]))))
    return MitchellSchaeffer2D, sims


def serving_stream(n: int = 64, seed: int = 7, rate: float = 200.0, side_range=(128, 1024),
                   horizon_range=(500.0, 2000.0)):
    """(request specs) as dicts: arrival, model key, BatchSim."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for k in range(n):
        model = "tp06" if rng.random() < 0.5 else "aliev_panfilov"
        side = int(rng.integers(side_range[0], side_range[1] + 1))
        p = float(rng.uniform(0.0, 0.4))
        ang = float(rng.uniform(0.0, 180.0))
        cross = bool(rng.random() < 0.5)
        horizon = float(rng.uniform(*horizon_range))
        tissue = fibrotic_tissue(side, p, 10_000 + k, fibre_deg=ang)
        n_ = side + 2
        if model == "tp06":
            s1 = S.StimCurrentCoord2D(0, 100, 1.0, 0, n_, 0, 6)
            s2 = S.StimCurrentCoord2D(310, 100, 1.0, n_ // 2, n_, 0, n_ // 2)
            thr = -40.0
        else:
            s1 = S.StimVoltageCoord2D(0, 1.0, 0, n_, 0, 6)
            s2 = S.StimVoltageCoord2D(32, 1.0, n_ // 2, n_, 0, n_ // 2)
            thr = 0.5
        stims = [s1, s2] if cross else [s1]
        mid = n_ // 2
        sim = BatchSim(tissue=tissue, t_max=horizon, stims=stims,
                       probe=_tissue_cell(tissue.mesh, mid, mid), act_threshold=thr,
                       act_start=0.0, apd_threshold=thr, cv_row=mid,
                       tag=dict(model=model, side=side, p=p, angle=ang, cross=cross,
                                horizon=horizon))
        out.append(dict(arrival=float(arrivals[k]), model=model, sim=sim))
    return {"tp06": TP062D, "aliev_panfilov": AlievPanfilov2D}, out
