"""Serving: a request loop with continuous batching on one GPU.

Requests arrive over wall-clock time.  A request is admitted when it has
arrived, a batch slot is free and its HBM footprint fits the budget; it then
joins the running batch of its model at the next launch boundary, so work
admitted later overlaps on the device with work already running.  Each model
has one ``TissueBatch``; the loop advances the batches in turn by short
chunks of steps, retires finished tissues (their activation map, probe trace
and APD90 are read back) and records each request's latency from arrival to
results.

A precision policy (``precision_for``) may route a request to a batch of
another precision: ``fp64_for_fibrosis`` serves the physiological models'
tissues whose fibrotic fraction reaches a threshold on the FP64 path, where
conduction close to failure lets float rounding move activation times by
more than a step.

A request ends in one of three states: ``done`` (results read back),
``cancelled`` (``Server.cancel(rid)`` or its ``deadline_s`` passed: it
leaves the queue, or its tissue leaves its batch at the next launch
boundary) or ``diverged`` (its potential held a NaN or an infinity at a
check, every ``check_every`` chunks, one launch per batch: the tissue is
dropped and the rest of the batch runs on).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .batch import BatchSim, TissueBatch


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of serving
    model: str              # key of the server's model classes
    sim: BatchSim
    precision: str | None = None      # set by the server: "fp32" or "fp64"
    footprint: int = 0
    admitted: float | None = None
    finished: float | None = None
    deadline_s: float | None = None   # cancelled when unfinished this long after arrival
    status: str = "pending"           # pending, active, done, cancelled, diverged
    error: str | None = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def fibrotic_fraction(sim: BatchSim) -> float:
    """Fraction of the tissue's interior nodes that do not conduct."""
    inner = np.asarray(sim.tissue.mesh)[1:-1, 1:-1]
    return float(np.count_nonzero(inner != 1)) / max(1, inner.size)


def fp64_for_fibrosis(models=("tp06",), threshold: float = 0.3, base: str = "fp32"):
    """A precision policy: requests of `models` whose tissue's fibrotic
    fraction is at least `threshold` run on the FP64 path, the rest on `base`."""
    models = tuple(models)

    def policy(req: "Request") -> str:
        if req.model in models and fibrotic_fraction(req.sim) >= threshold:
            return "fp64"
        return base
    return policy


class Server:
    def __init__(self, model_classes: dict, precision: str = "fp32", dt: float = 0.01,
                 dr: float = 0.25, hbm_budget: int | None = None, max_tissues: int = 256,
                 chunk_steps: int = 200, use_graph: bool = True, check_every: int = 1,
                 precision_for=None):
        import torch
        self.model_classes = dict(model_classes)
        self.precision = precision
        self.precision_for = precision_for
        self._batch_args = dict(dt=dt, dr=dr, max_tissues=max_tissues, use_graph=use_graph,
                                chunk_cap=chunk_steps)
        # one batch per (model, precision), made when a request first needs it
        self.batches: dict = {}
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.6 * free)
        self.hbm_budget = hbm_budget
        self.in_use = 0
        self.chunk_steps = chunk_steps
        self.completed: list[Request] = []
        self.cancelled: list[Request] = []
        self.diverged: list[Request] = []
        self.check_every = check_every
        self._cancel: set = set()
        self.peak_in_use = 0
        self.max_concurrent = 0

    def cancel(self, rid: int) -> None:
        """Cancels request `rid`: it is not admitted, or leaves its batch at
        the next launch boundary; results are not read."""
        self._cancel.add(rid)

    def batch(self, req: Request) -> TissueBatch:
        """The batch `req` runs in (its model and precision)."""
        if req.precision is None:
            req.precision = (self.precision_for(req) if self.precision_for
                             else self.precision)
        key = (req.model, req.precision)
        if key not in self.batches:
            self.batches[key] = TissueBatch(self.model_classes[req.model],
                                            precision=req.precision, **self._batch_args)
        return self.batches[key]

    def footprint(self, req: Request) -> int:
        b = self.batch(req)
        steps = int(np.ceil(req.sim.t_max / b.dt))
        return b.footprint(np.shape(req.sim.tissue.mesh), steps)

    def serve(self, requests: list, realtime: bool = True, log=None) -> dict:
        """Serve `requests` (sorted by arrival).  With realtime the arrival
        times are honoured against the wall clock; otherwise every request is
        available at once."""
        import torch
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            r.footprint = self.footprint(r)
            if r.footprint > self.hbm_budget:
                raise ValueError(f"request {r.rid} needs {r.footprint} B, above the budget")
        active: dict = {}
        t0 = time.perf_counter()
        chunks = 0

        def expired(r, now):
            return r.rid in self._cancel or (r.deadline_s is not None and
                                             now - r.arrival > r.deadline_s)

        def end(r, status, error=None):
            r.finished = time.perf_counter() - t0
            r.status, r.error = status, error
            (self.cancelled if status == "cancelled" else self.diverged).append(r)
            if log:
                log(r)

        while pending or active:
            now = time.perf_counter() - t0
            # cancellations: queued requests leave the queue, running ones
            # their batch
            for r in [r for r in pending if (not realtime or r.arrival <= now) and
                      expired(r, now)]:
                pending.remove(r)
                end(r, "cancelled")
            for key, r in list(active.items()):
                if expired(r, now):
                    self.batch(r).drop(r.sim.slot)
                    del active[key]
                    self.in_use -= r.footprint
                    end(r, "cancelled")
            # admission
            while pending and (not realtime or pending[0].arrival <= now):
                r = pending[0]
                b = self.batch(r)
                if self.in_use + r.footprint > self.hbm_budget or not b.free:
                    break
                pending.popleft()
                b.admit(r.sim)
                r.status = "active"
                r.admitted = time.perf_counter() - t0
                active[id(r.sim)] = r
                self.in_use += r.footprint
                self.peak_in_use = max(self.peak_in_use, self.in_use)
            self.max_concurrent = max(self.max_concurrent, len(active))
            if not active:
                if pending and realtime:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                continue
            # one chunk on every batch with work
            chunks += 1
            for b in self.batches.values():
                if not b.sims:
                    continue
                for sim in b.advance(self.chunk_steps):
                    r = active.pop(id(sim))
                    r.finished = time.perf_counter() - t0
                    r.status = "done"
                    self.in_use -= r.footprint
                    self.completed.append(r)
                    if log:
                        log(r)
                # divergence: a check enqueued behind this chunk, read when done
                for slot in b.poll_nonfinite():
                    sim = b.drop(slot)
                    r = active.pop(id(sim))
                    self.in_use -= r.footprint
                    end(r, "diverged", f"non-finite potential by global step {b.step}")
                if self.check_every and chunks % self.check_every == 0:
                    b.check_nonfinite()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out = dict(completed=len(self.completed), cancelled=len(self.cancelled),
                   diverged=len(self.diverged), wall_s=wall, peak_hbm_bytes=self.peak_in_use,
                   max_concurrent=self.max_concurrent)
        if self.completed:
            lat = np.array([r.latency for r in self.completed])
            first = min(r.arrival for r in self.completed)
            last = max(r.finished for r in self.completed)
            out.update(sims_per_s=len(self.completed) / max(1e-9, last - first),
                       p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)))
        return out

    def close(self):
        for b in self.batches.values():
            b.close()
