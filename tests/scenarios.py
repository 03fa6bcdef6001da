"""Shared scenario builders for the parity tests: the same tissue, model and
initial perturbation on the GPU path (finitewave_amd) and on the oracle."""
from __future__ import annotations

import numpy as np

MODEL_CLASSES = {
    "aliev_panfilov": ("AlievPanfilov2D", "AlievPanfilov3D"),
    "barkley": ("Barkley2D", "Barkley3D"),
    "mitchell_schaeffer": ("MitchellSchaeffer2D", "MitchellSchaeffer3D"),
    "fenton_karma": ("FentonKarma2D", "FentonKarma3D"),
    "bueno_orovio": ("BuenoOrovio2D", "BuenoOrovio3D"),
    "luo_rudy91": ("LuoRudy912D", "LuoRudy913D"),
    "tp06": ("TP062D", "TP063D"),
    "courtemanche": ("Courtemanche2D", "Courtemanche3D"),
}

# a potential that starts an upstroke in each model's units
EXCITE = {"aliev_panfilov": 1.0, "barkley": 1.0, "mitchell_schaeffer": 1.0,
          "fenton_karma": 1.0, "bueno_orovio": 1.5, "luo_rudy91": 10.0, "tp06": 10.0,
          "courtemanche": 10.0}


def boxed_mesh(shape, fibrosis=0.0, seed=0):
    """All-tissue mesh with a zero outer layer and optional i.i.d. code-2
    fibrosis (the Diffuse pattern's 1 + (random <= p))."""
    rng = np.random.RandomState(seed)
    mesh = np.ones(shape, dtype=np.int8)
    if fibrosis > 0:
        mesh = (1 + (rng.random_sample(shape) <= fibrosis)).astype(np.int8)
    sl = [slice(None)] * len(shape)
    for a in range(len(shape)):
        for edge in (0, -1):
            s = list(sl)
            s[a] = edge
            mesh[tuple(s)] = 0
    return mesh


def random_fibers(shape, seed=1):
    """Unit fibre vectors with random directions (the reference accepts any
    vectors; unit length keeps the explicit scheme inside its stability
    bound at dt = 0.01, dr = 0.25)."""
    rng = np.random.RandomState(seed)
    f = rng.normal(size=(*shape, len(shape)))
    return f / np.linalg.norm(f, axis=-1, keepdims=True)


def layered_fibers(shape, lo_deg=-60.0, hi_deg=60.0):
    """Fibre angle rotating linearly along the last axis, in the plane of the
    first two axes (the 3D slab configuration)."""
    nk = shape[-1]
    ang = np.deg2rad(lo_deg + (hi_deg - lo_deg) * np.clip(np.arange(nk) - 1, 0, nk - 3) /
                     max(1, nk - 3))
    f = np.zeros((*shape, 3))
    f[..., 0] = np.cos(ang)
    f[..., 1] = np.sin(ang)
    return f


def make_pair(name, dim, shape, stencil="iso", fibrosis=0.0, fibers=None, conductivity=1.0,
              dt=0.01, dr=0.25, precision="fp64", special=None, excite_slab=3):
    """(product model, oracle sim), both initialised, with the same excited
    slab at low i; the product model has no engine yet."""
    import finitewave_amd as fw
    from oracle.oracle import OracleSim

    mesh = boxed_mesh(shape, fibrosis)
    cls = getattr(fw, MODEL_CLASSES[name][dim - 2])
    tissue = (fw.CardiacTissue2D if dim == 2 else fw.CardiacTissue3D)(list(shape))
    tissue.mesh = mesh.copy()
    tissue.fibers = fibers
    tissue.conductivity = conductivity
    if special is not None:
        tissue.special_boundaries = special
    model = cls()
    model.dt, model.dr = dt, dr
    model.prog_bar = False
    model.precision = precision
    model.cardiac_tissue = tissue
    if stencil == "sym":
        model.stencil = fw.SymmetricStencil2D()
    model.initialize()
    sim = OracleSim(name, tissue.mesh, dt, dr, conductivity=conductivity, fibers=fibers,
                    special_boundaries=special, stencil=stencil)
    sim.initialize()
    sl = (slice(1, 1 + excite_slab),) + (slice(None),) * (dim - 1)
    exc = np.zeros(shape, dtype=bool)
    exc[sl] = True
    exc &= tissue.mesh == 1
    model.u[exc] = EXCITE[name]
    sim.u[exc] = EXCITE[name]
    return model, sim


def oracle_state(sim, names):
    return {n: sim.state[n] for n in names}
