"""CPU oracle for the finitewave stepping path -- TEST INFRASTRUCTURE ONLY.

A restatement of finitewave 0.8.5's step semantics over numpy arrays, with the
per-node sweeps in C (``fw_oracle.c``, built by ``oracle/Makefile``).  It is
the checker the HIP path is compared against and the CPU baseline that
bench.py times; the product package never imports it.

What it restates (paths relative to the finitewave source tree):

* ``core/model/cardiac_model.py:91-211``  -- initialize / compute_weights /
  run / check_termination: stim -> diffusion -> ionic -> trackers ->
  ``t += dt`` -> ``step += 1`` -> swap -> commands -> termination.
* ``cpuwave*/stencil/*.py`` numpy prologues (half-step diffusion with
  ``np.roll``) and epilogues (``w*D*dt/dr**2``, centre ``+= 1``); the njit
  loops are in C.
* every ionic model's parameter defaults, initial values and the
  ``u_new`` initialisation (``u_new = u.copy()`` only for LR91, TP06 and
  Courtemanche).
* ``core/stimulation``: ``t >= stim.t and not passed`` -> apply ->
  ``passed = t >= stim.t + duration``; voltage sets, current adds
  ``dt*I`` with an optional clamp, both on ``mask & (mesh == 1)``.
* ``ActivationTime2DTracker._track`` and point-probe trackers, gated by
  ``start_time <= t <= end_time`` and ``step % k == 0`` (``tracker.py:77-81``).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libfworacle.so"
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I8 = ctypes.POINTER(ctypes.c_int8)
_I64 = ctypes.POINTER(ctypes.c_int64)


def lib():
    """Load the oracle library (raises if ``make -C oracle`` has not run)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(str(_LIB_PATH))
        L.orc_weights_iso2d.argtypes = [_D, _I8, _D, _D, ctypes.c_int64, ctypes.c_int64]
        L.orc_weights_aniso2d.argtypes = [_D, _I8, _D, _D, _D, _D, ctypes.c_int64, ctypes.c_int64]
        L.orc_weights_sym2d.argtypes = [_D, _I8, _D, _D, _D, _D, ctypes.c_int64, ctypes.c_int64]
        L.orc_weights_iso3d.argtypes = [_D, _I8, _D, _D, _D, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_int64]
        L.orc_weights_aniso3d.argtypes = [_D, _I8, ctypes.POINTER(_D), ctypes.c_int64,
                                          ctypes.c_int64, ctypes.c_int64]
        for name in ("orc_diffusion_iso2d", "orc_diffusion_aniso2d"):
            getattr(L, name).argtypes = [_D, _D, _D, _I64, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_int64]
        for name in ("orc_diffusion_iso3d", "orc_diffusion_aniso3d"):
            getattr(L, name).argtypes = [_D, _D, _D, _I64, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_int64, ctypes.c_int64]
        L.orc_ionic.argtypes = [ctypes.c_int, _D, _D, ctypes.POINTER(_D), _I64, ctypes.c_int64,
                                ctypes.c_double, _D]
        L.orc_ionic.restype = ctypes.c_int
        L.orc_act_time.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_double, ctypes.c_double]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray, ty=_D):
    return a.ctypes.data_as(ty)


# ---------------------------------------------------------------------------
# Model tables: parameter order is the reference kernel's argument order after
# ``dt``; state order is the order of the arrays the model's run_ionic_kernel
# passes (for Courtemanche that passes ``self.xr, self.xs`` into the kernel's
# ``xs, xr`` slots, courtemanche_2d.py:149 vs :569).
# ---------------------------------------------------------------------------

def _crn_params():
    Vc = 20100
    Vj = Vc * 0.68
    Vup = Vj * 0.06 * 0.92
    Vrel = Vj * 0.06 * 0.08
    return [("gna", 7.8), ("gnab", 0.000674), ("gk1", 0.09), ("gkr", 0.0294),
            ("gks", 0.129), ("gto", 0.1652), ("gcal", 0.1238), ("gcab", 0.00113),
            ("gkur_coeff", 1), ("F", 96485.0), ("T", 310.0), ("R", 8314.0),
            ("Vc", Vc), ("Vj", Vj), ("Vup", Vup), ("Vrel", Vrel), ("ibk", 0.0),
            ("cao", 1.8), ("nao", 140), ("ko", 5.4), ("caupmax", 15), ("kup", 0.00092),
            ("kmnai", 10), ("kmko", 1.5), ("kmnancx", 87.5), ("kmcancx", 1.38),
            ("ksatncx", 0.1), ("kmcmdn", 0.00238), ("kmtrpn", 0.0005), ("kmcsqn", 0.8),
            ("trpnmax", 0.07), ("cmdnmax", 0.05), ("csqnmax", 10.0), ("inacamax", 1600),
            ("inakmax", 0.6), ("ipcamax", 0.275), ("krel", 30), ("iupmax", 0.005),
            ("kq10", 3)]


MODELS = {
    # aliev_panfilov_2d.py:74-103
    "aliev_panfilov": dict(
        id=0, D=1.0, copy_u_new=False,
        params=[("a", 0.1), ("k", 8.0), ("eap", 0.01), ("mu_1", 0.2), ("mu_2", 0.3)],
        state=[("v", 0.0)], init_u=0.0),
    # barkley_2d.py:55-82
    "barkley": dict(
        id=1, D=1.0, copy_u_new=False,
        params=[("a", 0.75), ("b", 0.02), ("eap", 0.02)],
        state=[("v", 0)], init_u=0.0),
    # mitchell_schaeffer_2d.py:40-69
    "mitchell_schaeffer": dict(
        id=2, D=1.0, copy_u_new=False,
        params=[("tau_close", 150), ("tau_open", 120), ("tau_in", 0.3), ("tau_out", 6),
                ("u_gate", 0.13)],
        state=[("h", 1.0)], init_u=0.0),
    # fenton_karma_2d.py:74-107
    "fenton_karma": dict(
        id=3, D=1.0, copy_u_new=False,
        params=[("tau_d", 0.172), ("tau_o", 12.5), ("tau_r", 130), ("tau_si", 127),
                ("tau_v_m", 18.2), ("tau_v_p", 10), ("tau_w_m", 80), ("tau_w_p", 1020),
                ("k", 10), ("u_c", 0.13), ("uc_si", 0.85)],
        state=[("v", 1.0), ("w", 1.0)], init_u=0.0),
    # bueno_orovio_2d.py:95-150
    "bueno_orovio": dict(
        id=4, D=1.0, copy_u_new=False,
        params=[("u_o", 0.0), ("u_u", 1.55), ("theta_v", 0.3), ("theta_w", 0.13),
                ("theta_v_m", 0.006), ("theta_o", 0.006), ("tau_v1_m", 60),
                ("tau_v2_m", 1150), ("tau_v_p", 1.4506), ("tau_w1_m", 60), ("tau_w2_m", 15),
                ("k_w_m", 65), ("u_w_m", 0.03), ("tau_w_p", 200), ("tau_fi", 0.11),
                ("tau_o1", 400), ("tau_o2", 6), ("tau_so1", 30.0181), ("tau_so2", 0.9957),
                ("k_so", 2.0458), ("u_so", 0.65), ("tau_s1", 2.7342), ("tau_s2", 16),
                ("k_s", 2.0994), ("u_s", 0.9087), ("tau_si", 1.8875), ("tau_w_inf", 0.07),
                ("w_inf_", 0.94)],
        state=[("v", 1.0), ("w", 1.0), ("s", 0.0)], init_u=0.0),
    # luo_rudy91_2d.py:52-124
    "luo_rudy91": dict(
        id=5, D=0.1, copy_u_new=True,
        params=[("gna", 23.0), ("gsi", 0.09), ("gk", 0.282), ("gk1", 0.6047), ("gkp", 0.0183),
                ("gb", 0.03921), ("ko", 5.4), ("ki", 145.0), ("nai", 18.0), ("nao", 140.0),
                ("cao", 1.8), ("R", 8.314), ("T", 310.0), ("F", 96.5), ("PR_NaK", 0.01833)],
        state=[("m", 0.0017), ("h", 0.9832), ("j", 0.995484), ("d", 0.000003), ("f", 1.0),
               ("x", 0.0057), ("cai", 0.0002)], init_u=-84.5),
    # tp06_2d.py:74-202
    "tp06": dict(
        id=6, D=0.154, copy_u_new=True,
        params=[("ko", 5.4), ("cao", 2.0), ("nao", 140.0), ("Vc", 0.016404), ("Vsr", 0.001094),
                ("Vss", 0.00005468), ("Bufc", 0.2), ("Kbufc", 0.001), ("Bufsr", 10.),
                ("Kbufsr", 0.3), ("Bufss", 0.4), ("Kbufss", 0.00025), ("Vmaxup", 0.006375),
                ("Kup", 0.00025), ("Vrel", 0.102), ("k1_", 0.15), ("k2_", 0.045),
                ("k3", 0.060), ("k4", 0.005), ("EC", 1.5), ("maxsr", 2.5), ("minsr", 1.),
                ("Vleak", 0.00036), ("Vxfer", 0.0038), ("R", 8314.472), ("F", 96485.3415),
                ("T", 310.0), ("RTONF", 26.71376), ("CAPACITANCE", 0.185), ("gkr", 0.153),
                ("pKNa", 0.03), ("gk1", 5.405), ("gna", 14.838), ("gbna", 0.00029),
                ("KmK", 1.0), ("KmNa", 40.0), ("knak", 2.724), ("gcal", 0.00003980),
                ("gbca", 0.000592), ("knaca", 1000), ("KmNai", 87.5), ("KmCa", 1.38),
                ("ksat", 0.1), ("n_", 0.35), ("gpca", 0.1238), ("KpCa", 0.0005),
                ("gpk", 0.0146), ("gto", 0.294), ("gks", 0.392)],
        state=[("cai", 0.00007), ("casr", 1.3), ("cass", 0.00007), ("nai", 7.67),
               ("Ki", 138.3), ("m", 0.), ("h", 0.75), ("j", 0.75), ("xr1", 0.),
               ("xr2", 1.), ("xs", 0.), ("r", 0.), ("s", 1.), ("d", 0.), ("f", 1.),
               ("f2", 1.), ("fcass", 1.), ("rr", 1.), ("oo", 0.)], init_u=-84.5),
    # courtemanche_2d.py:27-140
    "courtemanche": dict(
        id=7, D=0.154, copy_u_new=True, params=_crn_params(),
        state=[("nai", 11.2), ("ki", 139), ("cai", 0.000102), ("caup", 1.6), ("carel", 1.1),
               ("m", 0.00291), ("h", 0.965), ("j_", 0.978), ("d", 0.000137),
               ("f", 0.999837), ("oa", 0.000592), ("oi", 0.9992), ("ua", 0.003519),
               ("ui", 0.9987), ("xr", 3.29e-05), ("xs", 0.0187), ("fca", 0.775),
               ("irel", 0), ("vrel", 1), ("urel", 0), ("wrel", 0.9)],
        init_u=-84.5),
}


# ---------------------------------------------------------------------------
# Stimuli and trackers (plain records)
# ---------------------------------------------------------------------------

@dataclass
class OStim:
    """kind 'voltage' or 'current'; region: tuple of slices or bool matrix.
    values: a voltage per application (StimVoltageListMatrix3D.stimulate,
    cpuwave3D/stimulation/stim_voltage_list_matrix_3d.py:60-72: the k-th
    application sets values[k])."""
    kind: str
    t: float
    value: float
    region: object
    duration: float = 0.0
    u_max: float | None = None
    passed: bool = False
    values: list | None = None
    count: int = 0

    def apply(self, sim: "OracleSim") -> None:
        if isinstance(self.region, np.ndarray):
            mask = (self.region > 0) & (sim.mesh == 1)
            if self.kind == "voltage":
                if self.values is not None:
                    sim.u[mask] = self.values[self.count]
                    self.count += 1
                    return
                sim.u[mask] = self.value
            else:
                sim.u[mask] += sim.dt * self.value
                if self.u_max is not None:
                    sim.u[mask] = np.where(sim.u[mask] > self.u_max, self.u_max, sim.u[mask])
            return
        roi = sim.u[self.region]
        mask = sim.mesh[self.region] == 1
        if self.kind == "voltage":
            roi[mask] = self.value
        else:
            roi[mask] += sim.dt * self.value
            if self.u_max is not None:
                v = roi[mask]
                roi[mask] = np.where(v > self.u_max, self.u_max, v)


@dataclass
class OTracker:
    """kind 'act_time' (threshold), 'probe' (cells, var), 'lat' (threshold;
    LocalActivationTime), 'period' (cells, threshold), 'ecg' (coords) or
    'spiral' (threshold; SpiralWaveCore)."""
    kind: str
    threshold: float = -40.0
    cells: list = field(default_factory=list)
    var: str = "u"
    start_time: float = 0.0
    end_time: float = math.inf
    step: int = 1
    data: object = None
    coords: object = None


@dataclass
class OCommand:
    """Runs ``fn(sim)`` once, after the first step that ends with
    ``t >= t`` (command.py:42-52, command_sequence.py:51-59)."""
    t: float
    fn: Callable
    passed: bool = False


class OracleSim:
    """CardiacModel restated on the CPU (cardiac_model.py:91-211)."""

    def __init__(self, model: str, mesh: np.ndarray, dt: float, dr: float,
                 conductivity=1.0, fibers=None, special_boundaries=None,
                 stencil: str | None = None, D_al=1, D_ac=1 / 9,
                 params: dict | None = None, init: dict | None = None,
                 D_model: float | None = None):
        spec = MODELS[model]
        self.model = model
        self.spec = spec
        self.mesh = np.array(mesh)
        self.dim = self.mesh.ndim
        self.dt = dt
        self.dr = dr
        self.conductivity = conductivity
        self.fibers = fibers
        self.special_boundaries = special_boundaries
        self.stencil = stencil or ("iso" if fibers is None else "aniso")
        self.D_al = D_al
        self.D_ac = D_ac
        self.D_model = spec["D"] if D_model is None else D_model
        self.params = dict(spec["params"])
        if params:
            self.params.update(params)
        self.init = dict(spec["state"])
        self.init_u = spec["init_u"]
        if init:
            self.init_u = init.get("u", self.init_u)
            self.init.update({k: v for k, v in init.items() if k != "u"})
        self.t = 0
        self.step = 0
        self.stims: list[OStim] = []
        self.trackers: list[OTracker] = []
        self.commands: list[OCommand] = []

    # -- initialize (cardiac_model.py:91-117 + subclass initialize) --
    def initialize(self) -> None:
        shape = self.mesh.shape
        self.u = np.zeros(shape, dtype=np.float64)
        self.u_new = self.u.copy()
        self.step = 0
        self.t = 0
        self.compute_weights()
        self.u = self.init_u * np.ones(shape, dtype=np.float64)
        if self.spec["copy_u_new"]:
            self.u_new = self.u.copy()
        self.state = {name: val * np.ones(shape, dtype=np.float64)
                      for name, val in self.init.items()}
        for s in self.stims:
            s.passed = False
        for cmd in self.commands:
            cmd.passed = False
        for tr in self.trackers:
            if tr.kind == "act_time":
                tr.data = -np.ones_like(self.u)
            elif tr.kind in ("lat", "period"):
                tr.data = None
            elif tr.kind == "spiral":
                tr.data = {"u_prev": self.u.copy(), "tips": []}
            else:
                tr.data = []

    # -- weights (stencil compute_weights, numpy prologue + C loop + epilogue) --
    def compute_weights(self) -> None:
        mesh = self.mesh
        if self.special_boundaries is not None:
            self.idx = np.flatnonzero((mesh == 1) & (self.special_boundaries == 0)).astype(np.int64)
        else:
            self.idx = np.flatnonzero(mesh == 1).astype(np.int64)
        m = mesh.copy()
        m[m != 1] = 0
        m8 = np.ascontiguousarray(m.astype(np.int8))
        cond = self.conductivity * np.ones_like(m, dtype=np.float64)
        L = lib()
        shape = mesh.shape
        dr2 = self.dr ** 2
        if self.stencil == "iso":
            D = np.zeros((self.dim, *shape))
            for a in range(self.dim):
                D[a] = 0.5 * (cond + np.roll(cond, -1, axis=a))
            K = 5 if self.dim == 2 else 7
            w = np.zeros((*shape, K), dtype=np.float64)
            if self.dim == 2:
                L.orc_weights_iso2d(_p(w), _p(m8, _I8), _p(D[0]), _p(D[1]), *shape)
                centre = 2
            else:
                L.orc_weights_iso3d(_p(w), _p(m8, _I8), _p(D[0]), _p(D[1]), _p(D[2]), *shape)
                centre = 3
            w = w * self.D_model * self.dt / dr2
            w[..., centre] += 1
        elif self.stencil == "aniso":
            if self.fibers is None:
                raise ValueError("Fibers must be provided for anisotropic diffusion.")
            f = np.asarray(self.fibers, dtype=np.float64)
            d = []
            for axis in range(self.dim):
                D = np.zeros((self.dim, *shape))
                for i in range(self.dim):
                    D[i] = (self.D_ac * (axis == i) +
                            (self.D_al - self.D_ac) * f[..., axis] * f[..., i])
                    D[i] = 0.5 * (D[i] * cond + np.roll(D[i], -1, axis=axis) *
                                  np.roll(cond, -1, axis=axis))
                d.extend(np.ascontiguousarray(D[i]) for i in range(self.dim))
            if self.dim == 2:
                w = np.zeros((*shape, 9))
                L.orc_weights_aniso2d(_p(w), _p(m8, _I8), _p(d[0]), _p(d[1]), _p(d[2]),
                                      _p(d[3]), *shape)
            else:
                w = np.zeros((*shape, 19))
                arr = (_D * 9)(*[_p(x) for x in d])
                L.orc_weights_aniso3d(_p(w), _p(m8, _I8), arr, *shape)
            self._keep = d
            w = w * self.D_model * self.dt / dr2
            w[..., 4] += 1
        elif self.stencil == "sym":
            if self.dim != 2 or self.fibers is None:
                raise ValueError("symmetric stencil is 2D and needs fibers")
            f = np.asarray(self.fibers, dtype=np.float64)
            D = np.zeros((4, *shape))
            for i in range(4):
                i0, i1 = i // 2, i % 2
                D[i] = self.D_ac * (i0 == i1) + (self.D_al - self.D_ac) * f[..., i0] * f[..., i1]
                D[i] *= self.conductivity
                D[i] = 0.25 * (D[i] + np.roll(D[i], -1, axis=0) + np.roll(D[i], -1, axis=1) +
                               np.roll(np.roll(D[i], -1, axis=0), -1, axis=1))
            w = np.zeros((*shape, 9))
            L.orc_weights_sym2d(_p(w), _p(m8, _I8), _p(D[0]), _p(D[1]), _p(D[2]), _p(D[3]),
                                *shape)
            w *= self.D_model * self.dt / self.dr ** 2
            w[..., 4] += 1
        else:
            raise ValueError(self.stencil)
        self.weights = np.ascontiguousarray(w)

    def _diffuse(self) -> None:
        L = lib()
        K = self.weights.shape[-1]
        fn = {(2, 5): L.orc_diffusion_iso2d, (2, 9): L.orc_diffusion_aniso2d,
              (3, 7): L.orc_diffusion_iso3d, (3, 19): L.orc_diffusion_aniso3d}[(self.dim, K)]
        fn(_p(self.u_new), _p(self.u), _p(self.weights), _p(self.idx, _I64),
           len(self.idx), *self.mesh.shape)

    def _ecg(self, coords) -> np.ndarray:
        """ECG2DTracker.calc_ecg / compute_ecg (ecg_2d_tracker.py:61-151,
        ecg_3d_tracker.py:100-140): one diffusion sweep of the pre-step u
        into a zero array, then per measuring point the sum over active nodes
        of (u_tr - u) / (d * dr), d the squared index distance, summed in
        index order."""
        L = lib()
        K = self.weights.shape[-1]
        fn = {(2, 5): L.orc_diffusion_iso2d, (2, 9): L.orc_diffusion_aniso2d,
              (3, 7): L.orc_diffusion_iso3d, (3, 19): L.orc_diffusion_aniso3d}[(self.dim, K)]
        u_tr = np.zeros_like(self.u)
        fn(_p(u_tr), _p(self.u), _p(self.weights), _p(self.idx, _I64),
           len(self.idx), *self.mesh.shape)
        shape = self.mesh.shape
        co = np.unravel_index(self.idx, shape)
        diff = (u_tr.reshape(-1)[self.idx] - self.u.reshape(-1)[self.idx])
        out = np.zeros(len(coords))
        for e, (x, y, z) in enumerate(np.atleast_2d(coords)):
            if self.dim == 2:
                d = (x - co[0]) ** 2 + (y - co[1]) ** 2 + z ** 2
            else:
                d = (x - co[0]) ** 2 + (y - co[1]) ** 2 + (z - co[2]) ** 2
            d = d.astype(np.float64)
            ok = d > 0
            acc = 0.0
            for v in (diff[ok] / (d[ok] * self.dr)):
                acc += v
            out[e] = acc
        return out

    def _ionic_args(self):
        names = [n for n, _ in self.spec["state"]]
        key = tuple(self.state[n].ctypes.data for n in names)
        if getattr(self, "_ion_key", None) != key:
            self._ion_ptrs = (_D * len(names))(*[_p(self.state[n]) for n in names])
            self._ion_params = np.array([float(self.params[n]) for n, _ in self.spec["params"]])
            self._ion_key = key
        return self._ion_ptrs, self._ion_params

    def _ionic(self) -> None:
        ptrs, params = self._ionic_args()
        rc = lib().orc_ionic(self.spec["id"], _p(self.u_new), _p(self.u), ptrs,
                             _p(self.idx, _I64), len(self.idx), self.dt, _p(params))
        assert rc == 0

    def one_step(self) -> None:
        """Diffusion + ionic only (no stim/trackers/time)."""
        self._diffuse()
        self._ionic()

    # -- run (cardiac_model.py:130-189) --
    def run(self, t_max: float, initialize: bool = True,
            on_step: Callable | None = None) -> None:
        if initialize:
            self.initialize()
        if t_max < self.t:
            raise ValueError("t_max must be greater than current t.")
        iters = int(np.ceil((t_max - self.t) / self.dt))
        max_iters = int(np.ceil(t_max / self.dt))
        L = lib()
        for _ in range(iters):
            for s in self.stims:
                if self.t >= s.t and not s.passed:
                    s.apply(self)
                    s.passed = self.t >= s.t + s.duration
            self._diffuse()
            self._ionic()
            for tr in self.trackers:
                if tr.start_time > self.t or self.t > tr.end_time:
                    continue
                if self.step % tr.step != 0:
                    continue
                if tr.kind == "act_time":
                    L.orc_act_time(_p(tr.data), _p(self.u), self.u.size, float(tr.threshold),
                                   float(self.t))
                elif tr.kind in ("lat", "period"):
                    _lat_track(tr, self.u, self.t)
                elif tr.kind == "ecg":
                    tr.data.append(self._ecg(tr.coords))
                elif tr.kind == "spiral":
                    _spiral_track(tr, self.u, self.t, self.step)
                else:
                    src = self.u if tr.var == "u" else self.state[tr.var]
                    cells = tuple(np.atleast_2d(tr.cells).T)
                    tr.data.append(src[cells].copy())
            self.t += self.dt
            self.step += 1
            self.u_new, self.u = self.u, self.u_new
            for cmd in self.commands:
                if not cmd.passed and self.t >= cmd.t:
                    cmd.passed = True
                    cmd.fn(self)
            if on_step is not None:
                on_step(self)
            if (self.t > t_max) or (self.step > max_iters):
                break


def _lat_track(tr, u, t) -> None:
    """LocalActivationTime2DTracker._track / cross_threshold
    (local_activation_time_2d_tracker.py:60-102); kind 'period' restricts the
    layers to tr.cells (period_2d_tracker.py:45-61).  tr.data holds
    {'act_t': [layers], 'activated': bool array}."""
    if tr.data is None:
        n = len(np.atleast_2d(tr.cells)) if tr.kind == "period" else None
        tr.data = {"act_t": [-np.ones(n if n is not None else u.shape)],
                   "activated": np.zeros(u.shape, dtype=bool)}
    st = tr.data
    cross = (u >= tr.threshold) & (~st["activated"])
    st["activated"] = np.where(cross, True, st["activated"])
    back = (u < tr.threshold) & st["activated"]
    st["activated"] = np.where(back, False, st["activated"])
    if tr.kind == "period":
        cross = cross[tuple(np.atleast_2d(tr.cells).T)]
    if np.any(st["act_t"][-1][cross] > -1):
        st["act_t"].append(-np.ones(cross.shape))
    st["act_t"][-1] = np.where(cross, t, st["act_t"][-1])


def _tip_cross(u, i, j, thr) -> int:
    # _apply_threshold (spiral_wave_core_2d_tracker.py:175-208)
    if u[i, j] >= thr and (u[i + 1, j] < thr or u[i, j + 1] < thr or u[i + 1, j + 1] < thr):
        return 1
    if u[i, j] < thr and (u[i + 1, j] >= thr or u[i, j + 1] >= thr or u[i + 1, j + 1] >= thr):
        return 1
    return 0


def _tip_pos(i, j, u, un, thr):
    # _correct_tip_pos (spiral_wave_core_2d_tracker.py:108-172), same operation order
    AC = u[i, j] - u[i, j + 1] + u[i + 1, j + 1] - u[i + 1, j]
    GC = u[i, j + 1] - u[i, j]
    BC = u[i + 1, j] - u[i, j]
    DC = u[i, j] - thr
    AD = un[i, j] - un[i, j + 1] + un[i + 1, j + 1] - un[i + 1, j]
    GD = un[i, j + 1] - un[i, j]
    BD = un[i + 1, j] - un[i, j]
    DD = un[i, j] - thr
    Q = BC * AD - BD * AC
    R_ = GC * AD - GD * AC
    S = DC * AD - DD * AC
    QOnR = Q / R_
    SOnR = S / R_
    T = AC * QOnR
    U = AC * SOnR - BC + GC * QOnR
    V = GC * SOnR - DC
    disc = U * U - 4. * T * V
    if disc < 0:
        return None
    T2 = 2. * T
    if T2 == 0.:
        return None
    xn = (-U - math.sqrt(disc)) / T2
    xp = (-U + math.sqrt(disc)) / T2
    yn = -QOnR * xn - SOnR
    yp = -QOnR * xp - SOnR
    if 0 <= xn <= 1 and 0 <= yn <= 1:
        return (xn, yn)
    if 0 <= xp <= 1 and 0 <= yp <= 1:
        return (xp, yp)
    return None


def _tips_2d(u_prev, u, thr) -> list:
    # _track_tip_line (spiral_wave_core_2d_tracker.py:211-249): margin 5
    out = []
    ni, nj = u.shape
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(5, ni - 5):
            for j in range(5, nj - 5):
                c = _tip_cross(u_prev, i, j, thr)
                if c == 1:
                    c += _tip_cross(u, i, j, thr)
                if c == 2:
                    pos = _tip_pos(i, j, u_prev, u, thr)
                    if pos is not None:
                        out.append((j + pos[1], i + pos[0]))
    return out


def _spiral_track(tr, u, t, step) -> None:
    """SpiralWaveCore2DTracker._track (spiral_wave_core_2d_tracker.py:70-82)
    and its 3D per-slice form (spiral_wave_core_3d_tracker.py:18-31); the
    previous u starts as the initial u (initialize, :36-47).  tr.data holds
    {'u_prev': array, 'tips': [(x, y, (z,) t, step)]}."""
    st = tr.data
    if u.ndim == 2:
        for x, y in _tips_2d(st["u_prev"], u, tr.threshold):
            st["tips"].append((x, y, t, step))
    else:
        for k in range(u.shape[2]):
            for x, y in _tips_2d(st["u_prev"][:, :, k], u[:, :, k], tr.threshold):
                st["tips"].append((x, y, k, t, step))
    st["u_prev"] = u.copy()


def half_step_fields(model_D_al, model_D_ac, fibers, conductivity, shape, dim):
    """Exposed for tests: the anisotropic half-step fields in the order
    xx, xy, (xz,) yx, yy, (yz, zx, zy, zz)."""
    cond = conductivity * np.ones(shape, dtype=np.float64)
    f = np.asarray(fibers, dtype=np.float64)
    out = []
    for axis in range(dim):
        for i in range(dim):
            D = model_D_ac * (axis == i) + (model_D_al - model_D_ac) * f[..., axis] * f[..., i]
            out.append(0.5 * (D * cond + np.roll(D, -1, axis=axis) * np.roll(cond, -1, axis=axis)))
    return out


def time_fold(t0: float, dt: float, n: int) -> np.ndarray:
    """Model time before each of n steps, accumulated as the reference does
    (``t += dt``, cardiac_model.py:176)."""
    out = np.empty(n, dtype=np.float64)
    t = t0
    for k in range(n):
        out[k] = t
        t += dt
    return out


def set_threads(n: int | None = None) -> int:
    L = lib()
    if n:
        L.orc_set_num_threads(int(n))
    return L.orc_num_threads()
