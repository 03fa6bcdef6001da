"""fw_run_noalloc (no allocation) against fw_plan_create + fw_run, through the
C ABI: the same potentials, states and activation maps, bit for bit, on two
calls in a row, the second with a smaller tissue, for the plan shapes whose
field tables the entry point stages in its static scratch array (an FP32 3D
plan with fibres tabulated along z: double and float tables along k; an FP32
2D plan with conductivity tabulated along one axis: float copies) and for one
that stages nothing (FP64, uniform fields)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import finitewave_amd as fw
from tests.scenarios import boxed_mesh, layered_fibers


def _model(shape, precision, kind, seed):
    dim = len(shape)
    tissue = (fw.CardiacTissue2D if dim == 2 else fw.CardiacTissue3D)(list(shape))
    tissue.mesh = boxed_mesh(shape, 0.1, seed=seed)
    if kind == "layered":
        tissue.fibers = layered_fibers(shape, -50.0 - seed, 70.0)
    elif kind == "axis":
        cond = np.ones(shape)
        cond *= (0.6 + 0.4 * np.linspace(0, 1, shape[0]))[:, None]
        tissue.conductivity = cond
    elif kind == "uniform":
        f = np.zeros((*shape, dim))
        f[..., 0], f[..., 1] = np.cos(0.4), np.sin(0.4)
        tissue.fibers = f
    m = (fw.AlievPanfilov2D if dim == 2 else fw.AlievPanfilov3D)()
    m.dt, m.dr, m.prog_bar = 0.01, 0.25, False
    m.precision = precision
    m.cardiac_tissue = tissue
    m.initialize()
    exc = np.zeros(shape, dtype=bool)
    exc[1:4] = True
    m.u[exc & (tissue.mesh == 1)] = 1.0
    return m


def _run_both(model, steps):
    """The plan path and the no-allocation path from the same state: the
    potential after `steps` steps, the state and the activation map of each."""
    import torch
    from finitewave_amd import _native
    from finitewave_amd.engine import Engine, _stream
    L = _native.lib()
    out = []
    for noalloc in (False, True):
        eng = Engine(model)
        eng.upload()
        act = torch.full(model.u.shape, -1, dtype=torch.int32, device=eng.dev)
        track = _native.fw_track_desc()
        track.act = act.data_ptr()
        track.act_thr = 0.5
        if noalloc:
            rc = L.fw_run_noalloc(ctypes.byref(eng.plan.desc), eng.A.data_ptr(),
                                  eng.B.data_ptr(), eng._state_ptrs, 0, steps,
                                  ctypes.byref(track), _stream())
        else:
            rc = L.fw_run(eng.plan.handle, eng.A.data_ptr(), eng.B.data_ptr(),
                          eng._state_ptrs, 0, steps, ctypes.byref(track), _stream())
        _native.check(rc, "fw_run_noalloc" if noalloc else "fw_run")
        torch.cuda.synchronize()
        u = (eng.B if steps % 2 else eng.A).cpu().numpy()
        out.append((u, [s.cpu().numpy() for s in eng.state], act.cpu().numpy()))
        eng.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dim,precision,kind", [(3, "fp32", "layered"), (2, "fp32", "axis"),
                                                (2, "fp64", "uniform")])
def test_noalloc_equals_plan_path_on_two_calls(gpu, dim, precision, kind):
    shapes = [(22, 20, 18), (14, 12, 16)] if dim == 3 else [(60, 52), (34, 40)]
    for k, shape in enumerate(shapes):  # the second call is smaller
        model = _model(shape, precision, kind, seed=k)
        (u0, s0, a0), (u1, s1, a1) = _run_both(model, 41)
        assert np.array_equal(u0, u1), (shape, float(np.max(np.abs(u0 - u1))))
        assert all(np.array_equal(x, y) for x, y in zip(s0, s1)), shape
        assert np.array_equal(a0, a1), shape
        assert np.count_nonzero(a0 >= 0) > 0


@pytest.mark.gpu
def test_noalloc_refuses_tables_beyond_its_scratch(gpu):
    """An FP32 plan with a per-node conductivity needs a float copy of N
    values per field; beyond FW_NOALLOC_SCRATCH_BYTES the call is refused
    before any device work."""
    from finitewave_amd import _native
    from finitewave_amd.engine import Engine
    n = 1100  # 2 fields x 1100^2 floats = 9.7 MB > 4 MiB
    tissue = fw.CardiacTissue2D([n, n])
    tissue.conductivity = 0.5 + 0.5 * np.random.RandomState(1).random_sample((n, n))
    m = fw.AlievPanfilov2D()
    m.dt, m.dr, m.prog_bar, m.precision = 0.01, 0.25, False, "fp32"
    m.cardiac_tissue = tissue
    m.initialize()
    eng = Engine(m)
    eng.upload()
    L = _native.lib()
    rc = L.fw_run_noalloc(ctypes.byref(eng.plan.desc), eng.A.data_ptr(), eng.B.data_ptr(),
                          eng._state_ptrs, 0, 1, None, None)
    assert rc == -1  # FW_E_INVALID
    assert b"static scratch" in L.fw_last_error()
    eng.close()


def test_noalloc_checks_its_arguments_without_a_device():
    """Refusals that happen on the host, before any HIP call: a null
    description, a malformed one, and tables beyond the static scratch
    (fake non-null device addresses are never dereferenced)."""
    from ctypes import byref, c_void_p
    from finitewave_amd import _native
    L = _native.lib()
    st = (c_void_p * 1)(64)
    assert L.fw_run_noalloc(None, 64, 128, st, 0, 1, None, None) == -1
    d = _native.fw_plan_desc()
    d.ndim, d.model, d.stencil, d.precision = 2, 0, _native.STENCIL_ISO, _native.PREC_FP32
    d.shape[0], d.shape[1], d.shape[2] = 1100, 1100, 1
    d.code = 256
    d.n_fields = 2
    for q in range(2):
        d.fields[q].mode, d.fields[q].data = _native.FIELD_FULL, 512
    d.D_model, d.dt, d.dr2, d.n_params = 1.0, 0.01, 0.0625, 5
    rc = L.fw_run_noalloc(byref(d), 64, 128, st, 0, 1, None, None)
    assert rc == -1 and b"static scratch" in L.fw_last_error()
    d.n_params = 4
    assert L.fw_run_noalloc(byref(d), 64, 128, st, 0, 1, None, None) == -1
    assert b"parameters" in L.fw_last_error()
