"""Pins of the CPU oracle (no GPU).

The reference ships no golden vectors; its own known-answer tests are the
physiological windows of tests/test_models_2d.py / test_models_3d.py (APD and
amplitude per model) and tests/test_trackers_2d.py (conduction velocity from
the activation-time map).  They are reproduced here on the oracle.  The
stencils are pinned by closed forms (interior and Neumann edge weights),
invariants (weights sum to one, constant fields are fixed points, mirror and
transpose symmetry) and special cases (axis-aligned fibres have no cross
terms).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from oracle.oracle import MODELS, OracleSim, OStim, OTracker, time_fold
from tests.scenarios import boxed_mesh, random_fibers


def _apd(u, dt, thr, beat=3):
    # tests/test_models_2d.py:74-106
    up = np.where((u[:-1] < thr) & (u[1:] >= thr))[0]
    dn = np.where((u[:-1] > thr) & (u[1:] <= thr))[0]
    if len(up) <= beat or len(dn) == 0:
        return None
    s = up[beat]
    e = dn[dn > s]
    return None if len(e) == 0 else (e[0] - s) * dt


# (model, current, duration, t_calc, t_prebeats, threshold, APD window, amplitude checks)
# tests/test_models_2d.py:108-293
WINDOWS = [
    ("aliev_panfilov", 5, 0.5, 80, 60, 0.1, (20, 25), ("approx_max", 1.0, 0.1)),
    ("barkley", 5, 0.1, 80, 60, 0.1, (1, 4), ("approx_max", 1.0, 0.1)),
    ("mitchell_schaeffer", 5, 0.5, 1000, 1000, 0.1, (250, 350), ("approx_max", 0.95, 0.1)),
    ("fenton_karma", 5, 0.5, 1000, 1000, 0.1, (100, 200), ("approx_max", 1.0, 0.1)),
    ("bueno_orovio", 5, 0.5, 1000, 1000, 0.1, (200, 300), ("approx_max", 1.4, 0.1)),
    ("luo_rudy91", 100, 1, 1000, 1000, -70, (350, 400), ("gt_max", 20, -80)),
    ("tp06", 100, 1, 1000, 1000, -70, (280, 320), ("gt_max", 20, -80)),
    ("courtemanche", 100, 1, 1000, 1000, -70, (200, 300), ("gt_max", 10, -80)),
]


def run_reference_protocol(name, cur, dur, t_calc, t_pre, dim=2):
    """tests/test_models_2d.py:8-72: 5x3 (x3) tissue, four current pulses on
    the first two rows, u recorded at [3, 1] every step."""
    shape = (5, 3) if dim == 2 else (5, 3, 3)
    sim = OracleSim(name, boxed_mesh(shape), 0.01, 0.25)
    region = (slice(0, 2), slice(0, shape[1])) + ((slice(0, shape[2]),) if dim == 3 else ())
    for k in range(4):
        sim.stims.append(OStim("current", k * t_pre, cur, region, duration=dur))
    cell = [3, 1] if dim == 2 else [3, 1, 1]
    tr = OTracker("probe", cells=cell)
    sim.trackers.append(tr)
    sim.run(3 * t_pre + t_calc)
    return np.squeeze(np.array(tr.data))


@pytest.mark.parametrize("case", WINDOWS, ids=[w[0] for w in WINDOWS])
def test_reference_model_windows_2d(case):
    name, cur, dur, t_calc, t_pre, thr, (lo, hi), amp = case
    u = run_reference_protocol(name, cur, dur, t_calc, t_pre)
    apd = _apd(u, 0.01, thr)
    assert apd is not None and lo <= apd <= hi, apd
    if amp[0] == "approx_max":
        assert abs(np.max(u) - amp[1]) <= amp[2]
        assert abs(np.min(u) - 0.0) <= 0.01
    else:
        assert np.max(u) > amp[1] and np.min(u) < amp[2]


@pytest.mark.parametrize("name", ["aliev_panfilov", "barkley"])
def test_reference_model_windows_3d(name):
    # tests/test_models_3d.py: tighter max-amplitude tolerance (0.02)
    case = [w for w in WINDOWS if w[0] == name][0]
    _, cur, dur, t_calc, t_pre, thr, (lo, hi), amp = case
    u = run_reference_protocol(name, cur, dur, t_calc, t_pre, dim=3)
    apd = _apd(u, 0.01, thr)
    assert lo <= apd <= hi
    assert abs(np.max(u) - amp[1]) <= 0.1


def test_activation_time_conduction_velocity():
    # tests/test_trackers_2d.py:7-22, 123-146: AP cable 12x3, speed in [1.5, 2]
    sim = OracleSim("aliev_panfilov", boxed_mesh((12, 3)), 0.01, 0.25)
    sim.stims.append(OStim("current", 0, 5, (slice(0, 5), slice(0, 3)), duration=0.5))
    tr = OTracker("act_time", threshold=0.5)
    sim.trackers.append(tr)
    sim.run(3)
    speed = 5 * 0.25 / tr.data[10, 1]
    assert 1.5 <= speed <= 2


def test_resting_states_are_stable():
    # with no stimulus the physiological models stay near rest for 200 ms
    for name in ("luo_rudy91", "tp06", "courtemanche"):
        sim = OracleSim(name, boxed_mesh((5, 3)), 0.01, 0.25)
        tr = OTracker("probe", cells=[2, 1])
        sim.trackers.append(tr)
        sim.run(200)
        u = np.squeeze(np.array(tr.data))
        assert u.max() < -75 and u.min() > -95, (name, u.min(), u.max())
    # the two-variable models at their rest state do not move at all
    for name in ("aliev_panfilov", "barkley", "mitchell_schaeffer"):
        sim = OracleSim(name, boxed_mesh((6, 6)), 0.01, 0.25)
        sim.run(5)
        assert np.all(sim.u == 0.0)


# ---------------------------------------------------------------- stencils
def _weights(shape, stencil, fibers=None, cond=1.0, fibrosis=0.0, D_ac=1 / 9, dr=0.25):
    sim = OracleSim("aliev_panfilov", boxed_mesh(shape, fibrosis), 0.01, dr, fibers=fibers,
                    conductivity=cond, stencil=stencil, D_ac=D_ac)
    sim.compute_weights()
    return sim


@pytest.mark.parametrize("dim,stencil", [(2, "iso"), (3, "iso"), (2, "aniso"), (3, "aniso"),
                                         (2, "sym")])
def test_weights_sum_to_one_and_constant_fixed_point(dim, stencil):
    shape = (12, 13) if dim == 2 else (9, 10, 11)
    fib = random_fibers(shape) if stencil != "iso" else None
    sim = _weights(shape, stencil, fibers=fib, fibrosis=0.15)
    w = sim.weights.reshape(-1, sim.weights.shape[-1])[sim.idx]
    assert np.max(np.abs(w.sum(axis=1) - 1.0)) < 1e-13
    sim.u = np.full(shape, 0.7)
    sim.u_new = np.zeros(shape)
    sim._diffuse()
    assert np.max(np.abs(sim.u_new.reshape(-1)[sim.idx] - 0.7)) < 1e-13
    # fibrosis and empty nodes never receive weight
    K = sim.weights.shape[-1]
    assert np.all(sim.weights.reshape(-1, K)[np.flatnonzero(sim.mesh.reshape(-1) != 1)][:, [
        q for q in range(K) if q != (2 if (dim, stencil) == (2, "iso") else
                                     3 if (dim, stencil) == (3, "iso") else 4)]] == 0)


def test_iso_closed_forms():
    # interior weight r = D*dt/dr**2, centre 1-4r (2D) / 1-6r (3D); a node whose
    # opposite neighbour is outside the tissue gets 2r (Neumann ghost doubling,
    # isotropic_stencil_2d.py:158)
    r = 1.0 * 0.01 / 0.25 ** 2
    sim = _weights((7, 7), "iso")
    w = sim.weights
    assert np.allclose(w[3, 3], [r, r, 1 - 4 * r, r, r], rtol=0, atol=1e-15)
    assert w[1, 3, 4] == pytest.approx(2 * r) and w[1, 3, 0] == 0.0
    sim3 = _weights((7, 7, 7), "iso")
    assert np.allclose(sim3.weights[3, 3, 3], [r, r, r, 1 - 6 * r, r, r, r], atol=1e-15)


def test_aniso_axis_aligned_fibres_have_no_cross_terms():
    shape = (9, 9)
    f = np.zeros((*shape, 2))
    f[..., 0] = 1.0
    sim = _weights(shape, "aniso", fibers=f)
    s = 0.01 / 0.25 ** 2
    w = sim.weights[4, 4]
    assert np.all(w[[0, 2, 6, 8]] == 0.0)
    assert w[1] == pytest.approx(1.0 * s, rel=1e-15) and w[7] == pytest.approx(1.0 * s, rel=1e-15)
    assert w[3] == pytest.approx(s / 9, rel=1e-14) and w[5] == pytest.approx(s / 9, rel=1e-14)


def test_aniso_transpose_and_mirror_symmetry():
    # transposing the grid and swapping the fibre components transposes the
    # 3x3 weight block; mirroring i with f_x -> -f_x mirrors it
    shape = (11, 11)
    mesh = boxed_mesh(shape, 0.2, seed=5)
    f = random_fibers(shape, seed=7)
    a = OracleSim("aliev_panfilov", mesh, 0.01, 0.25, fibers=f)
    a.compute_weights()
    b = OracleSim("aliev_panfilov", mesh.T.copy(), 0.01, 0.25,
                  fibers=np.ascontiguousarray(f.transpose(1, 0, 2)[..., ::-1]))
    b.compute_weights()
    wa = a.weights.reshape(*shape, 3, 3)
    wb = b.weights.reshape(*shape, 3, 3)
    inner = (slice(1, -1), slice(1, -1))
    assert np.allclose(wa.transpose(1, 0, 3, 2)[inner], wb[inner], atol=1e-15)
    fm = f[::-1].copy()
    fm[..., 0] *= -1
    c = OracleSim("aliev_panfilov", mesh[::-1].copy(), 0.01, 0.25, fibers=fm)
    c.compute_weights()
    wc = c.weights.reshape(*shape, 3, 3)
    assert np.allclose(wa[::-1, :, ::-1, :][inner], wc[inner], atol=1e-15)


# asymmetric_stencil_3d.py:26-44: offset (di, dj, dk) of each of the 19 weights
_W19 = [(-1, -1, 0), (-1, 0, 0), (-1, 1, 0), (0, -1, 0), (0, 0, 0), (0, 1, 0), (1, -1, 0),
        (1, 0, 0), (1, 1, 0), (0, -1, -1), (0, -1, 1), (0, 0, -1), (0, 0, 1), (0, 1, -1),
        (0, 1, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (1, 0, 1)]


def _block19(w):
    """(*shape, 19) weights as (*shape, 3, 3, 3) neighbourhood blocks."""
    out = np.zeros(w.shape[:3] + (3, 3, 3))
    for q, (a, b, c) in enumerate(_W19):
        out[..., a + 1, b + 1, c + 1] = w[..., q]
    return out


@pytest.mark.parametrize("perm", [(1, 0, 2), (2, 1, 0), (0, 2, 1)])
def test_aniso3d_axis_swap_symmetry(perm):
    """Exchanging two grid axes and the same two fibre components exchanges
    the same two axes of every node's 19-point block
    (asymmetric_stencil_3d.py:163-458), cross terms involving z included."""
    shape = (9, 10, 11)
    mesh = boxed_mesh(shape, 0.15, seed=4)
    f = random_fibers(shape, seed=8)
    a = OracleSim("aliev_panfilov", mesh, 0.01, 0.25, fibers=f, stencil="aniso")
    a.compute_weights()
    b = OracleSim("aliev_panfilov", np.ascontiguousarray(mesh.transpose(perm)), 0.01, 0.25,
                  fibers=np.ascontiguousarray(f.transpose(*perm, 3)[..., list(perm)]),
                  stencil="aniso")
    b.compute_weights()
    wa = _block19(a.weights.reshape(*shape, 19)).transpose(*perm, *(3 + np.array(perm)))
    wb = _block19(b.weights.reshape(*b.mesh.shape, 19))
    inner = (slice(1, -1),) * 3
    assert np.abs(wa[inner]).max() > 0.01
    assert np.allclose(wa[inner], wb[inner], atol=1e-15, rtol=0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_aniso3d_mirror_symmetry(axis):
    """Mirroring the grid along one axis and negating that fibre component
    mirrors every node's 19-point block along that axis."""
    shape = (9, 10, 11)
    mesh = boxed_mesh(shape, 0.15, seed=6)
    f = random_fibers(shape, seed=9)
    a = OracleSim("aliev_panfilov", mesh, 0.01, 0.25, fibers=f, stencil="aniso")
    a.compute_weights()
    fm = np.flip(f, axis).copy()
    fm[..., axis] *= -1
    b = OracleSim("aliev_panfilov", np.ascontiguousarray(np.flip(mesh, axis)), 0.01, 0.25,
                  fibers=fm, stencil="aniso")
    b.compute_weights()
    wa = np.flip(np.flip(_block19(a.weights.reshape(*shape, 19)), axis), 3 + axis)
    wb = _block19(b.weights.reshape(*shape, 19))
    inner = (slice(1, -1),) * 3
    assert np.allclose(wa[inner], wb[inner], atol=1e-15, rtol=0)


def test_aniso3d_reduces_to_2d_in_a_plane():
    # fibres in the xy-plane and a slab one tissue node thick in z: the
    # 19-point weights restricted to the plane equal the 9-point weights
    f3 = random_fibers((10, 11, 3), seed=2)
    f3[..., 2] = 0.0
    f3[:, :, 0] = f3[:, :, 1]
    f3[:, :, 2] = f3[:, :, 1]
    m3 = boxed_mesh((10, 11, 3))
    s3 = OracleSim("aliev_panfilov", m3, 0.01, 0.25, fibers=f3)
    s3.compute_weights()
    s2 = OracleSim("aliev_panfilov", m3[:, :, 1].copy(), 0.01, 0.25,
                   fibers=np.ascontiguousarray(f3[:, :, 1, :2]))
    s2.compute_weights()
    w3 = s3.weights[:, :, 1, :9]
    assert np.allclose(w3[1:-1, 1:-1], s2.weights[1:-1, 1:-1], atol=1e-15)
    assert np.all(s3.weights[:, :, 1, 9:] == 0.0)


# ---------------------------------------------------------------- time and stimuli
def test_time_fold_values():
    # cardiac_model.py:176 accumulates t += dt
    t = time_fold(0.0, 0.01, 31001)
    assert t[31000] == 309.99999999986306
    assert time_fold(0.0, 0.01, 1001)[1000] == 9.999999999999831


def test_current_stimulus_fencepost():
    # a 0.5 ms current at dt 0.01 is applied 51 times (stim.py:56)
    sim = OracleSim("aliev_panfilov", boxed_mesh((5, 5)), 0.01, 0.25)
    stim = OStim("current", 0.0, 1.0, (slice(2, 3), slice(2, 3)), duration=0.5)
    count = {"n": 0}
    orig = stim.apply

    def counting(s):
        count["n"] += 1
        orig(s)

    stim.apply = counting
    sim.stims.append(stim)
    sim.run(2.0)
    assert count["n"] == 51


GOLDEN = Path(__file__).resolve().parent / "golden"
ORACLE_NAMES = {"AlievPanfilov": "aliev_panfilov", "Barkley": "barkley",
                "MitchellSchaeffer": "mitchell_schaeffer", "FentonKarma": "fenton_karma",
                "BuenoOrovio": "bueno_orovio", "LuoRudy91": "luo_rudy91", "TP06": "tp06",
                "Courtemanche": "courtemanche"}


def _reference_tables():
    # written by tests/golden/extract_reference_tables.py from the reference's
    # source text (ast, nothing imported)
    return json.loads((GOLDEN / "reference_tables.json").read_text())


def test_reference_tables_cover_every_model_class():
    tables = _reference_tables()
    assert sorted(tables) == sorted(f"{k}{d}" for k in ORACLE_NAMES for d in ("2D", "3D"))
    for cls, t in tables.items():
        # the kernel's own parameter list: u_new, u, the state slots,
        # indexes, dt, the scalars -- the call fills every slot
        assert t["kernel_params"][:2] == ["u_new", "u"], cls
        assert len(t["kernel_params"]) == 4 + len(t["state_args"]) + len(t["scalars"]), cls
        assert t["state_vars"] == ["u"] + sorted(t["state_args"], key=t["state_vars"].index), cls
    # Courtemanche passes self.xr into the kernel's xs slot and vice versa
    # (courtemanche_2d.py:149 vs :569; courtemanche_3d.py:57 vs :90)
    for d in ("2D", "3D"):
        t = tables[f"Courtemanche{d}"]
        k = t["state_args"].index("xr")
        assert t["state_slots"][k] == "xs" and t["state_args"][k + 1] == "xs"
        assert t["state_slots"][k + 1] == "xr"


@pytest.mark.parametrize("cls_name", sorted(f"{k}{d}" for k in ORACLE_NAMES for d in ("2D", "3D")))
def test_oracle_and_product_tables_equal_the_reference(cls_name):
    """Parameters (in the kernel's argument order), state arrays (in the
    call's order), initial values, D_model and the u_new initialisation of
    the oracle and of the product equal what the reference's source holds."""
    import finitewave_amd as fw
    t = _reference_tables()[cls_name]
    spec = MODELS[ORACLE_NAMES[cls_name[:-2]]]
    # oracle
    assert [float(v) for _, v in spec["params"]] == [float(v) for v in t["scalars"]]
    assert [n for n, _ in spec["state"]] == t["state_args"]
    assert [float(v) for _, v in spec["state"]] == [float(v) for v in t["init"]]
    assert spec["D"] == t["D_model"] and spec["init_u"] == t["init_u"]
    assert spec["copy_u_new"] == t["copy_u_new"]
    # product
    m = getattr(fw, cls_name)()
    assert m._param_values() == [float(v) for v in t["scalars"]]
    assert [getattr(m, a) for a in t["scalar_args"]] == t["scalars"]
    assert m._state_call_order() == t["state_args"]
    assert m.D_model == t["D_model"] and m.init_u == t["init_u"]
    assert m.state_vars == t["state_vars"]
    assert m.npfloat == t["npfloat"]
    assert type(m)._copy_u_new == t["copy_u_new"]
    for (attr, init_attr, _), want in zip(
            sorted(type(m)._state, key=lambda s: t["state_args"].index(s[0])), t["init"]):
        assert getattr(m, init_attr) == want, (cls_name, attr)


def _minor_flux(d, m0, u0, m1, u1, m2, u2, m3, u3, m4, u4, m5, u5):
    # the FP32 path's minor flux (finitewave_amd/csrc/fw_kernels.hpp minor_flux)
    mh = m2 + m3 + m4 + m5
    ml = m0 + m1 + m2 + m3
    if not (m2 and m3 and mh >= 3 and ml >= 3):
        return 0.0
    sh = u2 + u3 + m4 * u4 + m5 * u5
    sl = u2 + u3 + m0 * u0 + m1 * u1
    return d * (sh * (0.25 if mh == 4 else 1 / 3) - sl * (0.25 if ml == 4 else 1 / 3))


def _flux_form_2d(u, mesh, d, s):
    """The FP32 path's 9-point flux form, evaluated in float64: the sum of
    four major face fluxes and eight minor (cross) fluxes of neighbour means."""
    dxx, dxy, dyx, dyy = d
    out = u.copy()
    ni, nj = u.shape
    for i in range(1, ni - 1):
        for j in range(1, nj - 1):
            if mesh[i, j] != 1:
                continue
            M = lambda a, b: int(mesh[i + a, j + b] == 1)  # noqa: E731
            U = lambda a, b: u[i + a, j + b]  # noqa: E731
            uc = u[i, j]
            fl = dxx[i - 1, j] * M(-1, 0) * (U(-1, 0) - uc)
            fl -= _minor_flux(dxy[i - 1, j], M(-1, -1), U(-1, -1), M(0, -1), U(0, -1), M(-1, 0),
                              U(-1, 0), 1, uc, M(-1, 1), U(-1, 1), M(0, 1), U(0, 1))
            fl += dyy[i, j - 1] * M(0, -1) * (U(0, -1) - uc)
            fl -= _minor_flux(dyx[i, j - 1], M(-1, -1), U(-1, -1), M(-1, 0), U(-1, 0), M(0, -1),
                              U(0, -1), 1, uc, M(1, -1), U(1, -1), M(1, 0), U(1, 0))
            fl += dyy[i, j] * M(0, 1) * (U(0, 1) - uc)
            fl += _minor_flux(dyx[i, j], M(-1, 1), U(-1, 1), M(-1, 0), U(-1, 0), M(0, 1),
                              U(0, 1), 1, uc, M(1, 1), U(1, 1), M(1, 0), U(1, 0))
            fl += dxx[i, j] * M(1, 0) * (U(1, 0) - uc)
            fl += _minor_flux(dxy[i, j], M(1, -1), U(1, -1), M(0, -1), U(0, -1), M(1, 0),
                              U(1, 0), 1, uc, M(1, 1), U(1, 1), M(0, 1), U(0, 1))
            out[i, j] = uc + s * fl
    return out


@pytest.mark.parametrize("trial", range(6))
def test_fp32_flux_form_is_the_reference_weighted_sum(trial):
    """The FP32 path rewrites the 9-point weighted sum as face fluxes (no
    weights, no weight arrays).  Algebraically the two must agree on every
    mesh: evaluated in float64 on random fibrotic meshes, random fibre angles
    and random u, the flux form equals sum_k w_k u_k with the oracle's
    weights (asymmetric_stencil_2d.py:204-433) to rounding."""
    rng = np.random.RandomState(100 + trial)
    shape = (22, 25)
    mesh = boxed_mesh(shape, fibrosis=rng.uniform(0, 0.45), seed=trial)
    ang = rng.uniform(0, np.pi)
    f = np.zeros((*shape, 2))
    f[..., 0], f[..., 1] = np.cos(ang), np.sin(ang)
    sim = OracleSim("aliev_panfilov", mesh, 0.01, 0.25, fibers=f, stencil="aniso")
    sim.initialize()
    u = rng.uniform(-1, 1, size=shape)
    w = sim.weights
    ref = u.copy()
    for c in sim.idx:
        i, j = divmod(int(c), shape[1])
        nb = u[i - 1:i + 2, j - 1:j + 2].reshape(-1)
        ref[i, j] = sum(w[i, j, k] * nb[k] for k in range(9))
    got = _flux_form_2d(u, mesh, sim._keep, 0.01 / 0.25 ** 2)
    assert np.max(np.abs(got - ref)) <= 1e-14


# ---- the reference's tracker known-answer tests (tests/test_trackers_2d.py,
# tests/test_trackers_3d.py), run on the oracle's trackers ------------------

def _barkley_spiral(dim=2):
    # spiral_model fixture (test_trackers_2d.py:25-39, test_trackers_3d.py:28-44)
    shape = (100, 100) if dim == 2 else (100, 100, 3)
    mesh = boxed_mesh(shape)
    sim = OracleSim("barkley", mesh, 0.01, 0.25)
    box1 = (slice(0, 100), slice(0, 3)) + ((slice(0, 3),) if dim == 3 else ())
    box2 = (slice(0, 50), slice(0, 100)) + ((slice(0, 3),) if dim == 3 else ())
    sim.stims.append(OStim("voltage", 0, 1, box1))
    sim.stims.append(OStim("voltage", 5, 1, box2))
    return sim


def _period_mean(layers, t_cells):
    lats = np.array(layers).T
    per = [np.diff(row[row != -1]) for row in lats]
    return np.mean(np.array([np.mean(x) if len(x) > 0 else np.nan for x in per]))


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_period_tracker_window(dim):
    # test_spiral_wave_period_2d_tracker / _3d_tracker: mean period 3.5 +- 0.2
    sim = _barkley_spiral(dim)
    cells = [[80, 80], [20, 70], [40, 10], [25, 90]]
    if dim == 3:
        cells = [c + [1] for c in cells]
    tr = OTracker("period", threshold=0.5, cells=cells, start_time=10, step=10)
    sim.trackers.append(tr)
    sim.run(20)
    assert abs(_period_mean(tr.data["act_t"], cells) - 3.5) <= 0.2


def test_oracle_spiral_core_tracker_window():
    # test_spiral_wave_core_2d_tracker: tips within x 32..38, y 47..53
    sim = _barkley_spiral(2)
    tr = OTracker("spiral", threshold=0.5, start_time=12, step=10)
    sim.trackers.append(tr)
    sim.run(20)
    tips = np.array(tr.data["tips"])
    assert len(tips) > 0
    x, y = tips[:, 0], tips[:, 1]
    assert x.min() >= 32 and x.max() <= 38 and y.min() >= 47 and y.max() <= 53


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_ecg_tracker_window(dim):
    # planar_model + test_ecg_2d_tracker / test_ecg_3d_tracker
    shape = (50, 5) if dim == 2 else (50, 5, 3)
    sim = OracleSim("aliev_panfilov", boxed_mesh(shape), 0.0015, 0.25)
    box = (slice(0, 5), slice(0, 5)) + ((slice(0, 3),) if dim == 3 else ())
    sim.stims.append(OStim("current", 5, 5, box, duration=0.5))
    tr = OTracker("ecg", step=10, coords=np.array([[25, 2, 0 if dim == 2 else 1]]))
    sim.trackers.append(tr)
    sim.run(15)
    ecg = np.array(tr.data).T[0]
    assert ecg.max() > 0.001 and ecg.min() < -0.001 and np.argmax(ecg) > 100


def test_oracle_local_activation_time_two_waves():
    # test_local_activation_time_2d_tracker: two layers, both waves 1.5..2 node speed
    sim = OracleSim("aliev_panfilov", boxed_mesh((12, 3)), 0.01, 0.25)
    sim.stims.append(OStim("current", 0, 5, (slice(0, 5), slice(0, 3)), duration=0.5))
    sim.stims.append(OStim("voltage", 45, 1, (slice(0, 5), slice(0, 10))))
    tr = OTracker("lat", threshold=0.5)
    sim.trackers.append(tr)
    sim.run(50)
    lats = np.array(tr.data["act_t"])
    assert len(lats) == 2
    lat1, lat2 = lats[:, 10, 1]
    assert lat1 < lat2
    assert 1.5 <= 5 * 0.25 / lat1 <= 2
    assert 1.5 <= 5 * 0.25 / (lat2 - 45) <= 2
