"""Cardiac tissue: mesh codes, conductivity, fibers.

Mirrors ``core/tissue/cardiac_tissue.py``, ``cpuwave2D/tissue/cardiac_tissue_2d.py``
and ``cpuwave3D/tissue/cardiac_tissue_3d.py``: mesh codes 0 (empty),
1 (tissue), 2 (fibrosis); the outer layer of the mesh is forced to 0 on every
assignment; ``special_boundaries`` marks conducting nodes that are never
updated.
"""
from __future__ import annotations

import copy

import numpy as np

from . import _native


class CardiacTissue:
    """Base class (core/tissue/cardiac_tissue.py:6)."""

    def __init__(self):
        self.meta = {}
        self.special_boundaries = None
        self.myo_indexes = None

    @property
    def mesh(self) -> np.ndarray:
        return self._mesh

    @mesh.setter
    def mesh(self, mesh: np.ndarray) -> None:
        # cardiac_tissue.py:38-52
        if mesh.ndim != self.meta["dim"]:
            raise ValueError("Mesh dimension must match the tissue dimension.")
        self._mesh = mesh
        self.add_boundaries()

    def compute_myo_indexes(self) -> None:
        # cardiac_tissue.py:54-63
        if self.special_boundaries is not None:
            self.myo_indexes = np.flatnonzero((self.mesh == 1) & (self.special_boundaries == 0))
            return
        self.myo_indexes = np.flatnonzero(self.mesh == 1)

    def add_boundaries(self) -> None:
        raise NotImplementedError

    def add_pattern(self, fibro_pattern) -> None:
        fibro_pattern.apply(self)

    def clean(self) -> None:
        self.mesh[self.mesh == 2] = 1

    def clone(self):
        return copy.deepcopy(self)

    # -- device-side node codes -------------------------------------------
    def node_codes(self) -> np.ndarray:
        """uint8 per node: bit 0 = conducting (mesh == 1), bit 1 = special
        boundary (conducts, never updated)."""
        mesh = np.asarray(self.mesh)
        conduct = mesh == 1
        code = conduct.astype(np.uint8) * _native.CODE_CONDUCT
        if self.special_boundaries is not None:
            frozen = conduct & (np.asarray(self.special_boundaries) != 0)
            code |= frozen.astype(np.uint8) * _native.CODE_FROZEN
        return np.ascontiguousarray(code)

    def check_boundary(self) -> None:
        """The reference indexes neighbours without bounds checks; an updated
        node on the outer layer would read outside the grid.  Refuse it."""
        code = self.node_codes()
        updated = (code & 3) == 1
        dim = updated.ndim
        for ax in range(dim):
            first = np.take(updated, 0, axis=ax)
            last = np.take(updated, updated.shape[ax] - 1, axis=ax)
            if first.any() or last.any():
                raise ValueError("tissue nodes (mesh == 1) on the outer layer of the mesh are "
                                 "not supported: the outer layer must be 0")


class CardiacTissue2D(CardiacTissue):
    """cpuwave2D/tissue/cardiac_tissue_2d.py:7-45"""

    def __init__(self, shape):
        super().__init__()
        self.meta["dim"] = 2
        self.meta["shape"] = shape
        self.mesh = np.ones(shape, dtype=np.int8)
        self.conductivity = 1.0
        self.fibers = None

    def add_boundaries(self) -> None:
        self._mesh[0, :] = 0
        self._mesh[:, 0] = 0
        self._mesh[-1, :] = 0
        self._mesh[:, -1] = 0


class CardiacTissue3D(CardiacTissue):
    """cpuwave3D/tissue/cardiac_tissue_3d.py:7-46"""

    def __init__(self, shape):
        super().__init__()
        self.meta["dim"] = 3
        self.meta["shape"] = shape
        self.mesh = np.ones(shape, dtype=np.int8)
        self.conductivity = 1.0
        self.fibers = None

    def add_boundaries(self) -> None:
        self.mesh[0, :, :] = 0
        self.mesh[:, 0, :] = 0
        self.mesh[:, :, 0] = 0
        self.mesh[-1, :, :] = 0
        self.mesh[:, -1, :] = 0
        self.mesh[:, :, -1] = 0
