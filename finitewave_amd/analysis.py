"""Per-tissue results derived from traces and activation maps.

APD90 follows the definition the reference's S1-S2 restitution protocol uses
(Tutorials/RestitutionCurve.ipynb): last upstroke through a threshold, peak
and minimum after it, duration until the trace falls below
``peak - 0.9 * (peak - min)``.  The conduction velocity is the distance over
the difference of activation times between two nodes of one row, as the
reference's tracker tests estimate it (tests/test_trackers_2d.py:145:
``5 * dr / act_t``).
"""
from __future__ import annotations

import numpy as np


def apd90(trace: np.ndarray, dt: float, threshold_up: float):
    u = np.asarray(trace, dtype=np.float64)
    up = np.where((u[:-1] < threshold_up) & (u[1:] >= threshold_up))[0]
    if len(up) == 0:
        return None
    s = up[-1]
    peak = np.max(u[s:])
    level = peak - 0.9 * (peak - np.min(u[s:]))
    idx = np.where(u[s:] < level)[0]
    if len(idx) == 0:
        return None
    return float(idx[0] * dt)


def conduction_velocity(act_t: np.ndarray, row: int, dr: float, frac=(0.25, 0.75)):
    """dr * (j2 - j1) / (t2 - t1) between the nodes at fractions `frac` of the
    row; None when either node never activated or the times coincide."""
    a = np.asarray(act_t)[row]
    n = len(a)
    j1, j2 = int(round(frac[0] * (n - 1))), int(round(frac[1] * (n - 1)))
    t1, t2 = a[j1], a[j2]
    if t1 < 0 or t2 < 0 or t2 == t1:
        return None
    return float(dr * (j2 - j1) / (t2 - t1))
