"""Per-tissue results derived from traces and activation maps.

APD90 follows the reference's S1-S2 restitution protocol
(Tutorials/RestitutionCurve.ipynb, cell 6): the trace of the S2 run only;
its last upstroke through a threshold; peak and minimum after it; the
duration until the trace falls below ``peak - 0.9 * (peak - min)``.  A trace
with no upstroke after S2 is a loss of capture (the notebook reports it and
stops the curve there); an upstroke that never repolarises gives no value.

Conduction velocity, two ways:
  * ``front_velocity``: Velocity2DCalculation.front_velocity
    (tools/velocity_2d_calculation.py:15-44) -- for every node activated
    last, its distance to the nearest node activated first, over the time
    between the two;
  * ``conduction_velocity``: the distance over the difference of activation
    times between two nodes of one row, as the reference's tracker tests
    estimate it (tests/test_trackers_2d.py:145: ``5 * dr / act_t``).
"""
from __future__ import annotations

import numpy as np


def apd90(trace: np.ndarray, dt: float, threshold_up: float, start: int = 0):
    """APD90 of the last upstroke of trace[start:] (RestitutionCurve.ipynb
    cell 6), in time units; None when there is no upstroke (loss of capture)
    or the upstroke does not repolarise below the 90 % level."""
    u = np.asarray(trace, dtype=np.float64)[start:]
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


def captured(trace: np.ndarray, threshold_up: float, start: int = 0) -> bool:
    """Whether trace[start:] holds an upstroke through threshold_up (the
    notebook's capture test before its APD90)."""
    u = np.asarray(trace, dtype=np.float64)[start:]
    return bool(np.any((u[:-1] < threshold_up) & (u[1:] >= threshold_up)))


def front_velocity(act_t: np.ndarray, dr: float):
    """Velocity2DCalculation.front_velocity (velocity_2d_calculation.py:
    15-44): the nodes with the earliest activation time (of those >= 0) and
    the nodes with the latest one; for each latest node its Euclidean index
    distance to the nearest earliest node, times dr, over the time between
    the two.  One value per latest node; None when nothing activated or all
    at once."""
    from scipy import spatial
    a = np.asarray(act_t, dtype=np.float64)
    if not np.any(a >= 0):
        return None
    t0 = np.min(a[a >= 0])
    t1 = np.max(a)
    if t1 == t0:
        return None
    start = np.argwhere(a == t0)
    last = np.argwhere(a == t1)
    dist, _ = spatial.KDTree(start).query(last)
    return dist * dr / (t1 - t0)


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
