"""Known-answer pins of the oracle's eight ionic restatements (CPU).

The oracle's C sweep (orc_ionic) is compared with a second restatement of each
reference kernel (tests/known_answers.py), written from the reference's
source independently of oracle/fw_oracle.c, with the parameters read from the
reference's source by ast (tests/golden/reference_tables.json).  States are
drawn over each model's physiological range so that every branch (gate
Heavisides, the u >= -40 mV switches, both signs of each current) is taken.

Tolerances:
  * Aliev-Panfilov, Barkley, Mitchell-Schaeffer: |oracle - exact| <= 1e-14
    with exact = rational arithmetic on the binary inputs (the FP64 result is
    a few roundings away; a changed constant moves it by ~1e-5 or more).
  * Fenton-Karma, Bueno-Orovio: the same, tanh evaluated by the host libm.
  * Luo-Rudy 91, TP06, Courtemanche: every output within 1e-12 relative (the
    two FP64 restatements share the operation order and the libm; only
    `x**n` may differ in its last bit), and every output's change over the
    step within 1e-6 relative plus four ulps of its old value.
Invariants on top: TP06's cai is never changed (tp06_2d.py:1098), the buffered
SR and subspace calcium satisfy their conservation equations exactly as the
closed-form quadratic defines them (tp06_2d.py:799-872), and gates at a
clamped potential follow an exact affine recursion (Rush-Larsen for TP06 and
Courtemanche, forward Euler for Luo-Rudy 91).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from oracle.oracle import _p, lib
from tests import known_answers as KA

GOLDEN = Path(__file__).resolve().parent / "golden"
TABLES = json.loads((GOLDEN / "reference_tables.json").read_text())
IDS = {"AlievPanfilov2D": 0, "Barkley2D": 1, "MitchellSchaeffer2D": 2, "FentonKarma2D": 3,
       "BuenoOrovio2D": 4, "LuoRudy912D": 5, "TP062D": 6, "Courtemanche2D": 7}
DT = 0.01


def _params(cls):
    t = TABLES[cls]
    return dict(zip(t["scalar_args"], t["scalars"])), np.array(t["scalars"], dtype=np.float64)


def _oracle_step(cls, u, states, u_new=None, L=None):
    """One oracle sweep over len(u) nodes; states: {call attribute: array};
    L: the oracle library (default: the built one).  Returns (u_new,
    {kernel slot name: array})."""
    import ctypes
    L = L or lib()
    t = TABLES[cls]
    n = len(u)
    u = np.ascontiguousarray(u, dtype=np.float64)
    un = np.zeros(n) if u_new is None else np.ascontiguousarray(u_new, dtype=np.float64).copy()
    arrs = [np.ascontiguousarray(states[a], dtype=np.float64).copy() for a in t["state_args"]]
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_p(a) for a in arrs])
    idx = np.arange(n, dtype=np.int64)
    _, par = _params(cls)
    rc = L.orc_ionic(IDS[cls], _p(un), _p(u), ptrs, _p(idx, ctypes.POINTER(ctypes.c_int64)),
                     n, DT, _p(par))
    assert rc == 0
    return un, dict(zip(t["state_slots"], arrs))


def _slots(cls, states):
    t = TABLES[cls]
    return {slot: states[a] for a, slot in zip(t["state_args"], t["state_slots"])}


# ---------------------------------------------------------------- light models
def _light_states(rng, n=400):
    u = np.concatenate([rng.uniform(-0.2, 1.3, n - 8),
                        [0.0, 0.125, 0.25, 0.5, 0.75, 1.0, 0.13, 0.3]])
    return u


def _abs(exact, got):
    d = abs(float(exact) - float(got))
    return d if d == d else np.inf  # NaN from the oracle is a mismatch


def light_deviation(cls, L=None):
    """Largest |oracle - exact| over the outputs of one step of a model
    without transcendental functions (or with tanh only)."""
    p, _ = _params(cls)
    if cls in ("AlievPanfilov2D", "Barkley2D", "MitchellSchaeffer2D"):
        rng = np.random.RandomState(11)
        u = _light_states(rng)
        var = "h" if cls.startswith("Mitchell") else "v"
        g = rng.uniform(0.0, 2.0, len(u)) if var == "v" else rng.uniform(0.0, 1.0, len(u))
        g[:4] = [0.0, 0.5, 1.0, 1.5] if var == "v" else [0.0, 0.25, 0.5, 1.0]
        un, st = _oracle_step(cls, u, {var: g}, L=L)
        fn = {"AlievPanfilov2D": KA.aliev_panfilov, "Barkley2D": KA.barkley,
              "MitchellSchaeffer2D": KA.mitchell_schaeffer}[cls]
        worst = 0.0
        for q in range(len(u)):
            du, o = fn(u[q], g[q], DT, p)
            worst = max(worst, _abs(du, un[q]), _abs(o[var], st[var][q]))
        return worst
    if cls == "FentonKarma2D":
        rng = np.random.RandomState(12)
        u = _light_states(rng)
        v, w = rng.uniform(0, 1, len(u)), rng.uniform(0, 1, len(u))
        un, st = _oracle_step(cls, u, {"v": v, "w": w}, L=L)
        worst = 0.0
        for q in range(len(u)):
            du, o = KA.fenton_karma(u[q], v[q], w[q], DT, p)
            worst = max(worst, _abs(du, un[q]), _abs(o["v"], st["v"][q]),
                        _abs(o["w"], st["w"][q]))
        return worst
    assert cls == "BuenoOrovio2D"
    rng = np.random.RandomState(13)
    u = np.concatenate([rng.uniform(-0.1, 1.6, 392), [0.0, 0.005, 0.006, 0.13, 0.3, 0.65,
                                                      0.9087, 1.55]])
    v, w, s_ = (rng.uniform(0, 1, len(u)) for _ in range(3))
    un, st = _oracle_step(cls, u, {"v": v, "w": w, "s": s_}, L=L)
    worst = 0.0
    for q in range(len(u)):
        du, o = KA.bueno_orovio(u[q], v[q], w[q], s_[q], DT, p)
        worst = max(worst, _abs(du, un[q]), *(_abs(o[k], st[k][q]) for k in ("v", "w", "s")))
    return worst


LIGHT_TOL = 1e-14


@pytest.mark.parametrize("cls", ["AlievPanfilov2D", "Barkley2D", "MitchellSchaeffer2D",
                                 "FentonKarma2D", "BuenoOrovio2D"])
def test_light_models_one_step_equals_exact_arithmetic(cls):
    worst = light_deviation(cls)
    assert worst <= LIGHT_TOL, worst


def test_fenton_karma_si_current_takes_the_fast_gate():
    # J_si takes the fast gate v, not w (fenton_karma_2d.py:326): with w
    # changed and v kept the potential's update does not move
    cls = "FentonKarma2D"
    rng = np.random.RandomState(12)
    u = _light_states(rng)
    v, w = rng.uniform(0, 1, len(u)), rng.uniform(0, 1, len(u))
    un, _ = _oracle_step(cls, u, {"v": v, "w": w})
    un2, _ = _oracle_step(cls, u, {"v": v, "w": np.zeros_like(w)})
    assert np.array_equal(un, un2)


# ---------------------------------------------------------------- physiological
def _tp06_states(rng, n):
    return dict(u=rng.uniform(-90, 45, n), cai=rng.uniform(5e-5, 8e-4, n),
                casr=rng.uniform(0.3, 4.0, n), cass=rng.uniform(5e-5, 5e-3, n),
                nai=rng.uniform(6, 12, n), Ki=rng.uniform(130, 140, n),
                m=rng.uniform(0, 1, n), h=rng.uniform(0, 1, n), j=rng.uniform(0, 1, n),
                xr1=rng.uniform(0, 1, n), xr2=rng.uniform(0, 1, n), xs=rng.uniform(0, 1, n),
                r=rng.uniform(0, 1, n), s=rng.uniform(0, 1, n), d=rng.uniform(0, 1, n),
                f=rng.uniform(0, 1, n), f2=rng.uniform(0, 1, n), fcass=rng.uniform(0, 1, n),
                rr=rng.uniform(0, 1, n), oo=rng.uniform(0, 1, n))


def _lr91_states(rng, n):
    return dict(u=rng.uniform(-95, 45, n), m=rng.uniform(0, 1, n), h=rng.uniform(0, 1, n),
                j=rng.uniform(0, 1, n), d=rng.uniform(0, 1, n), f=rng.uniform(0, 1, n),
                x=rng.uniform(0, 1, n), cai=rng.uniform(1e-4, 5e-3, n))


def _crn_states(rng, n):
    s = dict(u=rng.uniform(-90, 45, n), nai=rng.uniform(9, 14, n), ki=rng.uniform(130, 145, n),
             cai=rng.uniform(5e-5, 8e-4, n), caup=rng.uniform(0.5, 2.5, n),
             carel=rng.uniform(0.3, 2.0, n), irel=rng.uniform(0, 0.5, n))
    for g in ("m", "h", "j_", "d", "f", "oa", "oi", "ua", "ui", "xr", "xs", "fca", "vrel",
              "urel", "wrel"):
        s[g] = rng.uniform(0, 1, n)
    return s


PHYS = {"LuoRudy912D": (_lr91_states, KA.luo_rudy91),
        "TP062D": (_tp06_states, KA.tp06),
        "Courtemanche2D": (_crn_states, KA.courtemanche)}

# potentials at and beside each model's branch points, appended to the random
# states: the -40 mV switch of the sodium gates (all three), LR91's Xi branch
# at -100 mV, Courtemanche's removable singularity of a_m at -47.13 mV
EDGE_U = {"LuoRudy912D": [-40.2, -40.0, -39.8, -100.5, -100.0, -99.5, -104.0],
          "TP062D": [-40.2, -40.0, -39.8],
          "Courtemanche2D": [-40.2, -40.0, -39.8, -47.13]}


def _rel(a, b):
    d = abs(a - b) / max(abs(b), 1e-300)
    return d if d == d else np.inf  # a NaN on either side is a mismatch


def _incr_dev(old, got, want):
    """How far the oracle's increment got - old is from the restatement's
    want - old, in units of the allowed error: 1e-6 of the increment plus
    four units in the last place of the old value (the rounding of
    old + increment)."""
    d_got, d_want = got - old, want - old
    allow = 1e-6 * abs(d_want) + 4 * np.spacing(abs(old)) + 1e-300
    r = abs(d_got - d_want) / allow
    return r if r == r else np.inf


def phys_deviation(cls, L=None):
    """Worst deviation between the oracle and the second restatement over
    every output of one step from 300 random states, in units of the
    tolerance (pass <= 1): each new value within PHYS_TOL relative, and each
    increment within _incr_dev's allowance (slow concentrations such as
    Courtemanche's nai change by ~1e-13 relative per step, below any value
    tolerance)."""
    make, fn = PHYS[cls]
    rng = np.random.RandomState(21)
    edge = EDGE_U[cls]
    n = 300 + len(edge)
    s = make(rng, n)
    u = s.pop("u")
    u[-len(edge):] = edge
    un0 = rng.uniform(-1, 1, n)  # u_new holds the diffusion result already
    un, st = _oracle_step(cls, u, s, u_new=un0, L=L)
    p, _ = _params(cls)
    slots = _slots(cls, s)
    worst = 0.0
    for q in range(n):
        old = {k: float(v[q]) for k, v in slots.items()}
        du, out = fn(float(u[q]), old, DT, p)
        worst = max(worst, _rel(un[q], un0[q] + du) / PHYS_TOL,
                    _incr_dev(un0[q], un[q], un0[q] + du))
        for k, v in out.items():
            worst = max(worst, _rel(st[k][q], v) / PHYS_TOL, _incr_dev(old[k], st[k][q], v))
    return worst if np.isfinite(worst) else np.inf


PHYS_TOL = 1e-12


@pytest.mark.parametrize("cls", sorted(PHYS))
def test_physiological_models_one_step_equal_second_restatement(cls):
    worst = phys_deviation(cls)
    assert worst <= 1.0, worst


def test_tp06_cai_is_read_only_and_calcium_buffers_conserve():
    """cai, cai = calc_cai(...) keeps the old cai (tp06_2d.py:1098, Appendix
    quirk); casr and cass are the free concentrations whose free + buffered
    total changes by exactly dt times the net flux (tp06_2d.py:799-872)."""
    cls = "TP062D"
    rng = np.random.RandomState(22)
    n = 200
    s = _tp06_states(rng, n)
    u = s.pop("u")
    _, st = _oracle_step(cls, u, s)
    p, _ = _params(cls)
    assert np.array_equal(st["cai"], s["cai"])
    bufsr, kbufsr, bufss, kbufss = p["Bufsr"], p["Kbufsr"], p["Bufss"], p["Kbufss"]
    for q in range(n):
        sl = {k: float(v[q]) for k, v in _slots(cls, s).items()}
        _, out, tr = _tp06_fluxes(float(u[q]), sl, p)
        c0, c1 = sl["casr"], st["casr"][q]
        tot0 = c0 + bufsr * c0 / (c0 + kbufsr)
        tot1 = c1 + bufsr * c1 / (c1 + kbufsr)
        assert abs(tot1 - (tot0 + DT * (tr["iup"] - tr["irel"] - tr["ileak"]))) <= 1e-12 * tot0
        c0, c1 = sl["cass"], st["cass"][q]
        tot0 = c0 + bufss * c0 / (c0 + kbufss)
        tot1 = c1 + bufss * c1 / (c1 + kbufss)
        flux = (-tr["ixfer"] * (p["Vc"] / p["Vss"]) + tr["irel"] * (p["Vsr"] / p["Vss"]) +
                (-tr["ical"] * (1. / (2 * p["Vss"] * p["F"])) * p["CAPACITANCE"]))
        assert abs(tot1 - (tot0 + DT * flux)) <= 1e-11 * max(tot0, abs(DT * flux))


def _tp06_fluxes(u, sl, p):
    """The fluxes of the calcium balance from the second restatement."""
    import math
    du, out = KA.tp06(u, sl, DT, p)
    casr, cass, cai = sl["casr"], sl["cass"], sl["cai"]
    ileak = p["Vleak"] * (casr - cai)
    iup = p["Vmaxup"] / (1. + ((p["Kup"] * p["Kup"]) / (cai * cai)))
    ixfer = p["Vxfer"] * (cass - cai)
    irel = p["Vrel"] * out["oo"] * (casr - cass)
    # ical from the restatement's own update of d, f, f2, fcass
    F, R, T = p["F"], p["R"], p["T"]
    ical = p["gcal"] * out["d"] * out["f"] * out["f2"] * out["fcass"] * 4 * (u - 15) * \
        (F * F / (R * T)) * (0.25 * math.exp(2 * (u - 15) * F / (R * T)) * cass - p["cao"]) / \
        (math.exp(2 * (u - 15) * F / (R * T)) - 1.)
    return du, out, dict(ileak=ileak, iup=iup, ixfer=ixfer, irel=irel, ical=ical)


@pytest.mark.parametrize("cls,gates,kind", [
    ("TP062D", ["m", "h", "j_", "xr1", "xr2", "xs", "r", "s", "d", "f", "f2"], "rl"),
    ("Courtemanche2D", ["m", "h", "j_", "d", "f", "oa", "oi", "ua", "ui", "xr", "xs"], "rl"),
    ("LuoRudy912D", ["m", "h", "j_", "d", "f", "x"], "euler")])
def test_gates_at_a_clamped_potential_follow_their_recursion(cls, gates, kind):
    """With u held fixed a gate obeys g_{n+1} = g_inf + (g_n - g_inf) * E with
    E = exp(-dt/tau) (Rush-Larsen) or 1 - dt/tau (forward Euler), so the
    ratio of successive increments is the same on every step and equals
    the restatement's E."""
    make, fn = PHYS[cls]
    rng = np.random.RandomState(23)
    n = 24
    s = make(rng, n)
    u = s.pop("u")
    p, _ = _params(cls)
    hist = []
    cur = dict(s)
    for _ in range(4):
        _, st = _oracle_step(cls, u, cur)
        hist.append({k: v.copy() for k, v in st.items()})
        t = TABLES[cls]
        cur = {a: st[slot] for a, slot in zip(t["state_args"], t["state_slots"])}
    for g in gates:
        d1 = hist[1][g] - hist[0][g]
        d2 = hist[2][g] - hist[1][g]
        d3 = hist[3][g] - hist[2][g]
        ok = np.abs(d1) > 1e-9
        r1, r2 = d2[ok] / d1[ok], d3[ok] / d2[ok]
        assert np.allclose(r1, r2, rtol=1e-6, atol=1e-9), (cls, g)
        if kind == "rl":  # exp(-dt/tau) lies in (0, 1)
            assert np.all((r1 > 0) & (r1 < 1)), (cls, g)
        else:  # 1 - dt/tau: below 1, negative where tau < dt (Euler overshoot)
            assert np.all(r1 < 1), (cls, g)
