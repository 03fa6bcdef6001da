"""FP64 parity of the fused gfx950 kernel with the oracle, step by step.

Tolerance (FP64 path): per-step L-infinity of V <= 1e-12 over 100 steps and
every state variable within 1e-12 relative at the end.  Models without
transcendental functions (Aliev-Panfilov, Barkley, Mitchell-Schaeffer) must
match bit for bit: the kernel performs the reference's operations in the
reference's order without fused multiply-add.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.scenarios import boxed_mesh, layered_fibers, make_pair, random_fibers

MODELS = ["aliev_panfilov", "barkley", "mitchell_schaeffer", "fenton_karma", "bueno_orovio",
          "luo_rudy91", "tp06", "courtemanche"]
EXACT = {"aliev_panfilov", "barkley", "mitchell_schaeffer"}
V_TOL = 1e-12


def _run_compare(model, sim, steps):
    from finitewave_amd.engine import Engine
    eng = Engine(model)
    eng.upload()
    worst = 0.0
    for s in range(steps):
        eng.run_steps(s, 1, use_graph=False)
        sim.one_step()
        sim.u, sim.u_new = sim.u_new, sim.u
        u_gpu = eng.u_dev().cpu().numpy()
        assert np.isfinite(sim.u).all(), f"oracle diverged at step {s}"
        err = float(np.max(np.abs(u_gpu - sim.u)))
        worst = max(worst, err)
        assert err <= V_TOL, f"step {s}: max |dV| = {err:.3e}"
    eng.download()
    for name in eng.state_names:
        a = getattr(model, name)
        b = sim.state[name]
        scale = np.maximum(1.0, np.abs(b))
        rel = float(np.max(np.abs(a - b) / scale))
        assert rel <= 1e-12, f"state {name}: max rel diff {rel:.3e}"
    return worst


@pytest.mark.gpu
@pytest.mark.parametrize("name", MODELS)
def test_fp64_iso2d_all_models(gpu, name):
    # 40x36 with 10 % fibrosis and a per-node conductivity field (full mode)
    shape = (40, 36)
    cond = 0.5 + 0.5 * np.random.RandomState(3).random_sample(shape)
    model, sim = make_pair(name, 2, shape, fibrosis=0.1, conductivity=cond)
    worst = _run_compare(model, sim, 100)
    if name in EXACT:
        assert worst == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("name", MODELS)
def test_fp64_aniso2d_all_models(gpu, name):
    # per-node random fibres: the 9-point weights rebuilt from full fields
    shape = (36, 40)
    model, sim = make_pair(name, 2, shape, stencil="aniso", fibrosis=0.15,
                           fibers=random_fibers(shape))
    worst = _run_compare(model, sim, 100)
    if name in EXACT:
        assert worst == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["aliev_panfilov", "tp06", "courtemanche", "luo_rudy91"])
def test_fp64_iso3d(gpu, name):
    model, sim = make_pair(name, 3, (18, 16, 14), fibrosis=0.1)
    worst = _run_compare(model, sim, 100)
    if name in EXACT:
        assert worst == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["aliev_panfilov", "mitchell_schaeffer", "tp06", "bueno_orovio"])
def test_fp64_aniso3d_layered_fibers(gpu, name):
    # fibre angle rotating with k (per-axis field tables) and 20 % fibrosis
    shape = (16, 18, 20)
    model, sim = make_pair(name, 3, shape, stencil="aniso", fibrosis=0.2,
                           fibers=layered_fibers(shape))
    worst = _run_compare(model, sim, 100)
    if name in EXACT:
        assert worst == 0.0


@pytest.mark.gpu
def test_fp64_aniso3d_random_fibers(gpu):
    shape = (14, 15, 16)
    model, sim = make_pair("aliev_panfilov", 3, shape, stencil="aniso", fibrosis=0.1,
                           fibers=random_fibers(shape))
    assert _run_compare(model, sim, 100) == 0.0


@pytest.mark.gpu
def test_fp64_symmetric_stencil_weights_path(gpu):
    # SymmetricStencil2D runs through precomputed weights (FW_STENCIL_WEIGHTS)
    shape = (30, 32)
    model, sim = make_pair("aliev_panfilov", 2, shape, stencil="sym", fibrosis=0.1,
                           fibers=random_fibers(shape))
    assert _run_compare(model, sim, 100) == 0.0


@pytest.mark.gpu
def test_fp64_special_boundaries(gpu):
    shape = (30, 30)
    special = np.zeros(shape, dtype=np.int8)
    special[10:13, 5:25] = 1
    model, sim = make_pair("aliev_panfilov", 2, shape, special=special)
    assert _run_compare(model, sim, 100) == 0.0
