"""The ionic model classes with the reference's names, parameters, defaults
and initial values.

Each class mirrors the reference model of the same name
(``cpuwave2D/model/*.py``, ``cpuwave3D/model/*.py``): the parameters are
instance attributes, the state arrays are attributes named as in
``state_vars``, ``initialize()`` fills them with the ``init_*`` values, and
``select_stencil`` picks the isotropic stencil without fibres and the
asymmetric one with them.  The per-node update itself runs on the GPU.
"""
from __future__ import annotations

import numpy as np

from . import _native
from .model import CardiacModel
from .stencil import (AsymmetricStencil2D, AsymmetricStencil3D, IsotropicStencil2D,
                      IsotropicStencil3D)


class _IonicModel(CardiacModel):
    _name = ""
    _D = 1.0
    _params: list = []      # (attribute, default) in the kernel's argument order
    _state: list = []       # (attribute, init attribute, default) in state_vars order
    _call_order = None      # state attributes in the kernel's argument order
    _init_u = 0.0
    _copy_u_new = False     # u_new = u.copy() in initialize (LR91, TP06, Courtemanche)
    _dim = 2

    def __init__(self):
        super().__init__()
        self.D_model = self._D
        self.state_vars = ["u"] + [s for s, _, _ in self._state]
        self.npfloat = "float64"
        for name, value in self._params:
            setattr(self, name, value)
        self.init_u = self._init_u
        for name, init_name, value in self._state:
            setattr(self, name, np.ndarray)
            setattr(self, init_name, value)

    def initialize(self):
        super().initialize()
        shape = self.cardiac_tissue.mesh.shape
        self.u = self.init_u * np.ones(shape, dtype=self.npfloat)
        if self._copy_u_new:
            self.u_new = self.u.copy()
        for name, init_name, _ in self._state:
            setattr(self, name, getattr(self, init_name) * np.ones(shape, dtype=self.npfloat))

    def select_stencil(self, cardiac_tissue):
        if self._dim == 2:
            return IsotropicStencil2D() if cardiac_tissue.fibers is None else AsymmetricStencil2D()
        return IsotropicStencil3D() if cardiac_tissue.fibers is None else AsymmetricStencil3D()

    def _model_id(self) -> int:
        return _native.MODEL_IDS[self._name]

    def _param_values(self) -> list:
        return [float(getattr(self, name)) for name, _ in self._params]

    def _state_call_order(self) -> list:
        return list(self._call_order or [s for s, _, _ in self._state])


# ---- Aliev-Panfilov (aliev_panfilov_2d.py:74-103, aliev_panfilov_3d.py) ----
class AlievPanfilov2D(_IonicModel):
    _name = "aliev_panfilov"
    _params = [("a", 0.1), ("k", 8.0), ("eap", 0.01), ("mu_1", 0.2), ("mu_2", 0.3)]
    _state = [("v", "init_v", 0.0)]


class AlievPanfilov3D(AlievPanfilov2D):
    _dim = 3


# ---- Barkley (barkley_2d.py:55-82) ----
class Barkley2D(_IonicModel):
    _name = "barkley"
    _params = [("a", 0.75), ("b", 0.02), ("eap", 0.02)]
    _state = [("v", "init_v", 0)]


class Barkley3D(Barkley2D):
    _dim = 3


# ---- Mitchell-Schaeffer (mitchell_schaeffer_2d.py:40-69) ----
class MitchellSchaeffer2D(_IonicModel):
    _name = "mitchell_schaeffer"
    _params = [("tau_close", 150), ("tau_open", 120), ("tau_in", 0.3), ("tau_out", 6),
               ("u_gate", 0.13)]
    _state = [("h", "init_h", 1.0)]


class MitchellSchaeffer3D(MitchellSchaeffer2D):
    _dim = 3


# ---- Fenton-Karma (fenton_karma_2d.py:74-107) ----
class FentonKarma2D(_IonicModel):
    _name = "fenton_karma"
    _params = [("tau_d", 0.172), ("tau_o", 12.5), ("tau_r", 130), ("tau_si", 127),
               ("tau_v_m", 18.2), ("tau_v_p", 10), ("tau_w_m", 80), ("tau_w_p", 1020),
               ("k", 10), ("u_c", 0.13), ("uc_si", 0.85)]
    _state = [("v", "init_v", 1.0), ("w", "init_w", 1.0)]


class FentonKarma3D(FentonKarma2D):
    _dim = 3


# ---- Bueno-Orovio (bueno_orovio_2d.py:95-150) ----
class BuenoOrovio2D(_IonicModel):
    _name = "bueno_orovio"
    _params = [("u_o", 0.0), ("u_u", 1.55), ("theta_v", 0.3), ("theta_w", 0.13),
               ("theta_v_m", 0.006), ("theta_o", 0.006), ("tau_v1_m", 60), ("tau_v2_m", 1150),
               ("tau_v_p", 1.4506), ("tau_w1_m", 60), ("tau_w2_m", 15), ("k_w_m", 65),
               ("u_w_m", 0.03), ("tau_w_p", 200), ("tau_fi", 0.11), ("tau_o1", 400),
               ("tau_o2", 6), ("tau_so1", 30.0181), ("tau_so2", 0.9957), ("k_so", 2.0458),
               ("u_so", 0.65), ("tau_s1", 2.7342), ("tau_s2", 16), ("k_s", 2.0994),
               ("u_s", 0.9087), ("tau_si", 1.8875), ("tau_w_inf", 0.07), ("w_inf_", 0.94)]
    _state = [("v", "init_v", 1.0), ("w", "init_w", 1.0), ("s", "init_s", 0.0)]


class BuenoOrovio3D(BuenoOrovio2D):
    _dim = 3


# ---- Luo-Rudy 1991 (luo_rudy91_2d.py:52-124) ----
class LuoRudy912D(_IonicModel):
    _name = "luo_rudy91"
    _D = 0.1
    _init_u = -84.5
    _copy_u_new = True
    _params = [("gna", 23.0), ("gsi", 0.09), ("gk", 0.282), ("gk1", 0.6047), ("gkp", 0.0183),
               ("gb", 0.03921), ("ko", 5.4), ("ki", 145.0), ("nai", 18.0), ("nao", 140.0),
               ("cao", 1.8), ("R", 8.314), ("T", 310.0), ("F", 96.5), ("PR_NaK", 0.01833)]
    _state = [("m", "init_m", 0.0017), ("h", "init_h", 0.9832), ("j", "init_j", 0.995484),
              ("d", "init_d", 0.000003), ("f", "init_f", 1.0), ("x", "init_x", 0.0057),
              ("cai", "init_cai", 0.0002)]


class LuoRudy913D(LuoRudy912D):
    _dim = 3


# ---- TP06 epicardial (tp06_2d.py:74-202) ----
class TP062D(_IonicModel):
    _name = "tp06"
    _D = 0.154
    _init_u = -84.5
    _copy_u_new = True
    _params = [("ko", 5.4), ("cao", 2.0), ("nao", 140.0), ("Vc", 0.016404), ("Vsr", 0.001094),
               ("Vss", 0.00005468), ("Bufc", 0.2), ("Kbufc", 0.001), ("Bufsr", 10.),
               ("Kbufsr", 0.3), ("Bufss", 0.4), ("Kbufss", 0.00025), ("Vmaxup", 0.006375),
               ("Kup", 0.00025), ("Vrel", 0.102), ("k1_", 0.15), ("k2_", 0.045), ("k3", 0.060),
               ("k4", 0.005), ("EC", 1.5), ("maxsr", 2.5), ("minsr", 1.), ("Vleak", 0.00036),
               ("Vxfer", 0.0038), ("R", 8314.472), ("F", 96485.3415), ("T", 310.0),
               ("RTONF", 26.71376), ("CAPACITANCE", 0.185), ("gkr", 0.153), ("pKNa", 0.03),
               ("gk1", 5.405), ("gna", 14.838), ("gbna", 0.00029), ("KmK", 1.0),
               ("KmNa", 40.0), ("knak", 2.724), ("gcal", 0.00003980), ("gbca", 0.000592),
               ("knaca", 1000), ("KmNai", 87.5), ("KmCa", 1.38), ("ksat", 0.1), ("n_", 0.35),
               ("gpca", 0.1238), ("KpCa", 0.0005), ("gpk", 0.0146), ("gto", 0.294),
               ("gks", 0.392)]
    _state = [("cai", "init_cai", 0.00007), ("casr", "init_casr", 1.3),
              ("cass", "init_cass", 0.00007), ("nai", "init_nai", 7.67),
              ("Ki", "init_Ki", 138.3), ("m", "init_m", 0.), ("h", "init_h", 0.75),
              ("j", "init_j", 0.75), ("xr1", "init_xr1", 0.), ("xr2", "init_xr2", 1.),
              ("xs", "init_xs", 0.), ("r", "init_r", 0.), ("s", "init_s", 1.),
              ("d", "init_d", 0.), ("f", "init_f", 1.), ("f2", "init_f2", 1.),
              ("fcass", "init_fcass", 1.), ("rr", "init_rr", 1.), ("oo", "init_oo", 0.)]


class TP063D(TP062D):
    _dim = 3


# ---- Courtemanche (courtemanche_2d.py:27-140) ----
class Courtemanche2D(_IonicModel):
    _name = "courtemanche"
    _D = 0.154
    _init_u = -84.5
    _copy_u_new = True
    _params = [("gna", 7.8), ("gnab", 0.000674), ("gk1", 0.09), ("gkr", 0.0294),
               ("gks", 0.129), ("gto", 0.1652), ("gcal", 0.1238), ("gcab", 0.00113),
               ("gkur_coeff", 1), ("F", 96485.0), ("T", 310.0), ("R", 8314.0), ("Vc", 20100),
               ("Vj", None), ("Vup", None), ("Vrel", None), ("ibk", 0.0), ("cao", 1.8),
               ("nao", 140), ("ko", 5.4), ("caupmax", 15), ("kup", 0.00092), ("kmnai", 10),
               ("kmko", 1.5), ("kmnancx", 87.5), ("kmcancx", 1.38), ("ksatncx", 0.1),
               ("kmcmdn", 0.00238), ("kmtrpn", 0.0005), ("kmcsqn", 0.8), ("trpnmax", 0.07),
               ("cmdnmax", 0.05), ("csqnmax", 10.0), ("inacamax", 1600), ("inakmax", 0.6),
               ("ipcamax", 0.275), ("krel", 30), ("iupmax", 0.005), ("kq10", 3)]
    _state = [("nai", "init_nai", 11.2), ("ki", "init_ki", 139), ("cai", "init_cai", 0.000102),
              ("caup", "init_caup", 1.6), ("carel", "init_carel", 1.1), ("m", "init_m", 0.00291),
              ("h", "init_h", 0.965), ("j_", "init_j", 0.978), ("d", "init_d", 0.000137),
              ("f", "init_f", 0.999837), ("oa", "init_oa", 0.000592), ("oi", "init_oi", 0.9992),
              ("ua", "init_ua", 0.003519), ("ui", "init_ui", 0.9987), ("xr", "init_xr", 3.29e-05),
              ("xs", "init_xs", 0.0187), ("fca", "init_fca", 0.775), ("irel", "init_irel", 0),
              ("vrel", "init_vrel", 1), ("urel", "init_urel", 0), ("wrel", "init_wrel", 0.9)]
    # run_ionic_kernel passes self.xr, self.xs into the kernel's xs, xr slots
    # (courtemanche_2d.py:149 vs :569)
    _call_order = ["nai", "ki", "cai", "caup", "carel", "m", "h", "j_", "d", "f", "oa", "oi",
                   "ua", "ui", "xr", "xs", "fca", "irel", "vrel", "urel", "wrel"]

    def __init__(self):
        super().__init__()
        # courtemanche_2d.py:79-82
        self.Vj = self.Vc * 0.68
        self.Vup = self.Vj * 0.06 * 0.92
        self.Vrel = self.Vj * 0.06 * 0.08
        # the state_vars list of the reference (courtemanche_2d.py:28-29)
        self.state_vars = ["u", "nai", "ki", "cai", "caup", "carel", "m", "h", "j_", "d", "f",
                           "oa", "oi", "ua", "ui", "xr", "xs", "fca", "irel", "vrel", "urel",
                           "wrel"]


class Courtemanche3D(Courtemanche2D):
    _dim = 3
