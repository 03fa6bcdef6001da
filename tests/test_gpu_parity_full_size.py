"""FP64 parity at the benchmark configurations' real sizes.

The parity tests of test_gpu_parity.py run tissues of a few thousand nodes.
These run the kernel instances the benchmark runs, at their real sizes,
against the oracle:
  * TP06 2D 1026 x 1026 (1024^2 tissue), anisotropic 9-point stencil, uniform
    fibre angle 30 deg, dr 0.25 model units, dt 0.01, FP64: the uniform-field
    kernel instance with host-computed weights for nodes whose 3x3
    neighbourhood conducts.  100 steps from an excited band, per-step L-inf
    of V <= 1e-12 and every state within 1e-12 relative at the end.
  * Aliev-Panfilov 3D, a slab of the 258^3 configuration at its real
    cross-section: 34 x 258 x 258, fibre angle rotating from -60 to +60 deg
    along z, D_ac 0.111, diffuse fibrosis p = 0.2 drawn from RandomState(0)
    over the whole 258^3 grid (its first 34 planes, outer layer zeroed).  100
    FP64 steps bit for bit (Aliev-Panfilov's update is exact), the along-z
    field instance; then one FP32 step from the same state within 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest

import finitewave_amd as fw
from oracle.oracle import OracleSim


def _excite(model, sim, cols, value):
    sl = (slice(None), slice(1, 1 + cols)) if model.u.ndim == 2 else \
        (slice(1, 1 + cols), slice(None), slice(None))
    mask = np.zeros(model.u.shape, dtype=bool)
    mask[sl] = True
    mask &= model.cardiac_tissue.mesh == 1
    model.u[mask] = value
    sim.u[mask] = value


def _compare(model, sim, steps, tol):
    from finitewave_amd.engine import Engine
    eng = Engine(model)
    eng.upload()
    worst = 0.0
    for s in range(steps):
        eng.run_steps(s, 1, use_graph=False)
        sim.one_step()
        sim.u, sim.u_new = sim.u_new, sim.u
        err = float(np.max(np.abs(eng.u_dev().cpu().numpy() - sim.u)))
        worst = max(worst, err)
        assert err <= tol, f"step {s}: max |dV| = {err:.3e}"
    eng.download()
    for name in eng.state_names:
        a, b = getattr(model, name), sim.state[name]
        rel = float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))
        assert rel <= 1e-12, f"state {name}: {rel:.3e}"
    return worst, eng


@pytest.mark.gpu
def test_tp06_1026_aniso30_fp64_matches_oracle_for_100_steps(gpu):
    n = 1026
    tissue = fw.CardiacTissue2D([n, n])
    ang = np.deg2rad(30.0)
    f = np.zeros((n, n, 2))
    f[..., 0], f[..., 1] = np.cos(ang), np.sin(ang)
    tissue.fibers = f
    model = fw.TP062D()
    model.dt, model.dr, model.prog_bar = 0.01, 0.25, False
    model.precision = "fp64"
    model.cardiac_tissue = tissue
    model.initialize()
    sim = OracleSim("tp06", tissue.mesh, 0.01, 0.25, fibers=f)
    sim.initialize()
    _excite(model, sim, 10, 10.0)
    worst, eng = _compare(model, sim, 100, 1e-12)
    # every half-step field uniform: the plan runs the uniform-field instance
    d = eng.plan.desc
    assert d.n_fields == 4 and all(d.fields[q].mode == 0 for q in range(4))
    eng.close()


def _c3_slab(planes=34, side=258):
    rs = np.random.RandomState(0)
    mesh = (1 + (rs.random_sample((side,) * 3) <= 0.2)).astype(np.int8)[:planes].copy()
    for a in range(3):
        idx = [slice(None)] * 3
        for e in (0, -1):
            idx[a] = e
            mesh[tuple(idx)] = 0
    k = np.arange(side)
    phi = np.deg2rad(-60 + 120 * np.clip(k - 1, 0, side - 3) / (side - 3))
    f = np.zeros((planes, side, side, 3))
    f[..., 0] = np.cos(phi)[None, None, :]
    f[..., 1] = np.sin(phi)[None, None, :]
    return mesh, f


def _c3_model(mesh, f, precision):
    tissue = fw.CardiacTissue3D(list(mesh.shape))
    tissue.mesh = mesh.copy()
    tissue.fibers = f
    model = fw.AlievPanfilov3D()
    model.dt, model.dr, model.prog_bar = 0.01, 0.25, False
    model.precision = precision
    model.cardiac_tissue = tissue
    st = fw.AsymmetricStencil3D()
    st.D_al, st.D_ac = 1.0, 0.111
    model.stencil = st
    model.initialize()
    return model


@pytest.mark.gpu
def test_c3_slab_fp64_bitwise_and_fp32_one_step(gpu):
    mesh, f = _c3_slab()
    model = _c3_model(mesh, f, "fp64")
    sim = OracleSim("aliev_panfilov", mesh, 0.01, 0.25, fibers=f, stencil="aniso", D_ac=0.111)
    sim.initialize()
    _excite(model, sim, 5, 1.0)
    worst, eng = _compare(model, sim, 100, 0.0)
    assert worst == 0.0
    # fields uniform or tabulated along z: the along-z instance
    d = eng.plan.desc
    assert d.n_fields == 9
    assert all(d.fields[q].mode == 0 or (d.fields[q].mode == 1 and d.fields[q].axis == 2)
               for q in range(9))
    assert any(d.fields[q].mode == 1 for q in range(9))
    u64, v64 = eng.u_dev().cpu().numpy(), eng.state[0].cpu().numpy()
    eng.close()
    # one FP32 step from the FP64 state reached above
    m32 = _c3_model(mesh, f, "fp32")
    m32.u = u64.astype(np.float32)
    m32.u_new = np.zeros_like(m32.u)
    m32.v = v64.astype(np.float32)
    from finitewave_amd.engine import Engine
    e32 = Engine(m32)
    e32.upload()
    e32.run_steps(100, 1, use_graph=False)
    sim.u, sim.state["v"] = u64.copy(), v64.copy()
    sim.u_new = np.zeros_like(sim.u)
    sim.one_step()
    got = e32.u_dev().cpu().numpy().astype(np.float64)
    upd = sim.idx
    err = float(np.max(np.abs(got.reshape(-1)[upd] - sim.u_new.reshape(-1)[upd])))
    assert err <= 1e-5, err
    e32.close()
