"""Stencils: the half-step diffusion fields and the weight layout.

Mirrors ``core/stencil/stencil.py`` and the concrete stencils of
``cpuwave2D/stencil`` and ``cpuwave3D/stencil``.  The hot path never stores
the (*shape, K) weights: the fused kernel rebuilds them per node from the
node codes and the half-step fields below.  ``compute_weights`` still returns
the reference's array (built on the GPU) for callers that want it.
"""
from __future__ import annotations

import numpy as np

from . import _native


class Stencil:
    """Base class (core/stencil/stencil.py:4-46)."""

    kind = None  # _native.STENCIL_*
    K = None
    centre = None

    def compute_weights(self, model, cardiac_tissue) -> np.ndarray:
        """(*shape, K) weights scaled by D*dt/dr**2 with +1 on the centre,
        as the reference stencils return them."""
        from .engine import device_weights
        return device_weights(self, model, cardiac_tissue)

    def select_diffusion_kernel(self):
        from .engine import diffusion_kernel_for
        return diffusion_kernel_for(self)

    # half-step fields in the order the C ABI expects (fw_plan_desc.fields)
    def half_step_fields(self, model, cardiac_tissue) -> list:
        raise NotImplementedError

    @staticmethod
    def _mesh01(cardiac_tissue) -> np.ndarray:
        mesh = np.asarray(cardiac_tissue.mesh).copy()
        mesh[mesh != 1] = 0
        return mesh

    @staticmethod
    def _conductivity(cardiac_tissue, mesh) -> np.ndarray:
        return cardiac_tissue.conductivity * np.ones_like(mesh, dtype=np.float64)


class IsotropicStencil2D(Stencil):
    """5-point stencil (isotropic_stencil_2d.py:7-94).  Weight order:
    (i-1,j), (i,j-1), (i,j), (i,j+1), (i+1,j)."""

    kind = _native.STENCIL_ISO
    K = 5
    centre = 2
    ndim = 2

    def half_step_fields(self, model, cardiac_tissue) -> list:
        # compute_half_step_diffusion (isotropic_stencil_2d.py:71-94):
        # D[a] = 0.5 * (c + roll(c, -1, axis=a))
        mesh = self._mesh01(cardiac_tissue)
        c = self._conductivity(cardiac_tissue, mesh)
        return [0.5 * (c + np.roll(c, -1, axis=a)) for a in range(self.ndim)]


class IsotropicStencil3D(IsotropicStencil2D):
    """7-point stencil (isotropic_stencil_3d.py:10-74).  Weight order:
    (i-1), (j-1), (k-1), centre, (k+1), (j+1), (i+1)."""

    K = 7
    centre = 3
    ndim = 3


class AsymmetricStencil2D(Stencil):
    """9-point anisotropic stencil (asymmetric_stencil_2d.py:7-161).
    ``D_al`` / ``D_ac`` set the longitudinal / cross-fibre ratios."""

    kind = _native.STENCIL_ANISO
    K = 9
    centre = 4
    ndim = 2

    def __init__(self):
        self.D_al = 1
        self.D_ac = 1 / 9

    def compute_diffusion_components(self, fibers, ind0, ind1, D_al, D_ac):
        # asymmetric_stencil_2d.py:138-161
        return D_ac * (ind0 == ind1) + (D_al - D_ac) * fibers[..., ind0] * fibers[..., ind1]

    def half_step_fields(self, model, cardiac_tissue) -> list:
        fibers = cardiac_tissue.fibers
        if fibers is None:
            raise ValueError("Fibers must be provided for anisotropic diffusion.")
        mesh = self._mesh01(cardiac_tissue)
        c = self._conductivity(cardiac_tissue, mesh)
        f = np.asarray(fibers, dtype=np.float64)
        out = []
        # compute_half_step_diffusion (asymmetric_stencil_2d.py:97-136), per axis
        for axis in range(self.ndim):
            for comp in range(self.ndim):
                D = self.compute_diffusion_components(f, axis, comp, self.D_al, self.D_ac)
                D = D * np.ones(mesh.shape)
                out.append(0.5 * (D * c + np.roll(D, -1, axis=axis) * np.roll(c, -1, axis=axis)))
        return out


class AsymmetricStencil3D(AsymmetricStencil2D):
    """19-point anisotropic stencil (asymmetric_stencil_3d.py:11-105)."""

    K = 19
    centre = 4
    ndim = 3


class SymmetricStencil2D(AsymmetricStencil2D):
    """Flux-symmetric 9-point stencil (symmetric_stencil_2d.py:7-117).  It
    runs through the precomputed-weights path of the fused kernel."""

    kind = _native.STENCIL_SYM

    def half_step_fields(self, model, cardiac_tissue) -> list:
        fibers = cardiac_tissue.fibers
        if fibers is None:
            raise ValueError("Fibers must be provided for anisotropic diffusion.")
        mesh = self._mesh01(cardiac_tissue)
        f = np.asarray(fibers, dtype=np.float64)
        out = []
        # symmetric_stencil_2d.py:69-117: corner averages of D*conductivity
        for i in range(4):
            ind0, ind1 = i // 2, i % 2
            D = self.compute_diffusion_components(f, ind0, ind1, self.D_al, self.D_ac)
            D = D * np.ones(mesh.shape)
            D = D * cardiac_tissue.conductivity
            D = 0.25 * (D + np.roll(D, -1, axis=0) + np.roll(D, -1, axis=1) +
                        np.roll(np.roll(D, -1, axis=0), -1, axis=1))
            out.append(D)
        return out


def compress_field(arr: np.ndarray):
    """(mode, axis, value, data) for a half-step field: one value, a table
    along one axis, or the full array.  Compares bit patterns, so the
    compressed field reproduces every node's value exactly."""
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64))
    bits = a.view(np.uint64)
    if np.all(bits == bits.flat[0]):
        return _native.FIELD_UNIFORM, 0, float(a.flat[0]), None
    for ax in range(a.ndim):
        idx = [0] * a.ndim
        idx[ax] = slice(None)
        prof = bits[tuple(idx)]
        shape = [1] * a.ndim
        shape[ax] = -1
        if np.array_equal(bits, np.broadcast_to(prof.reshape(shape), bits.shape)):
            return _native.FIELD_AXIS, ax, 0.0, np.ascontiguousarray(a[tuple(idx)])
    return _native.FIELD_FULL, 0, 0.0, a
