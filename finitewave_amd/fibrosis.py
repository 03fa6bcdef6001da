"""Fibrosis patterns: Diffuse and Structural, 2D and 3D.

These build the mesh codes (1 tissue, 2 fibrosis) before a run, on the host.
They draw from the same generators the reference draws from and in the same
order: ``numpy.random.random`` over the whole grid for the diffuse pattern,
one ``random.random()`` per block for the structural one.  A user's seeded
meshes are therefore the reference's, bit for bit; a device generator could
not reproduce those streams, and the generation is one-off.

With ``device=True`` the pattern is drawn on the GPU instead
(``fw_fibrosis``): the same distribution, from a counter-based generator keyed
by ``seed`` and the node's (or block's) index, for large meshes where the host
loop is slow; it does not reproduce numpy's or Python's streams.

Reference: core/fibrosis/fibrosis_pattern.py, cpuwave2D/fibrosis/*.py,
cpuwave3D/fibrosis/*.py.
"""
from __future__ import annotations

import random

import numpy as np

__all__ = ["FibrosisPattern", "Diffuse2DPattern", "Structural2DPattern",
           "Diffuse3DPattern", "Structural3DPattern"]


class FibrosisPattern:
    """A pattern that produces mesh codes (core/fibrosis/fibrosis_pattern.py:4-51)."""

    def generate(self, shape=None, mesh=None, *, device=False, seed=0):
        """Returns mesh codes: a new all-tissue mesh of `shape` with the
        pattern in its region, or `mesh` itself with the pattern written
        into its region.  `shape` wins when both are given.  device: draw
        on the GPU, keyed by `seed`."""
        if shape is None and mesh is None:
            raise ValueError("Either shape or mesh must be provided.")
        if shape is not None:
            mesh = np.ones(shape, dtype=np.int8)
        if device:
            self._device_generate(mesh, seed)
            return mesh
        codes = self._codes(mesh.shape)
        region = self._region()
        mesh[region] = codes[region]
        return mesh

    def apply(self, cardiac_tissue, *, device=False, seed=0):
        """Writes the pattern into the tissue's mesh (through its setter)."""
        cardiac_tissue.mesh = self.generate(mesh=cardiac_tissue.mesh, device=device, seed=seed)

    def _blocks(self):
        return None  # one draw per node

    def _device_generate(self, mesh, seed):
        import ctypes

        from . import _native
        torch = _native.require_gpu()
        if mesh.dtype != np.int8 or mesh.ndim not in (2, 3):
            raise ValueError("device patterns take a 2D or 3D int8 mesh")
        lo, hi = [], []
        for sl, n in zip(self._region(), mesh.shape):
            a, b, step = sl.indices(n)
            if step != 1:
                raise ValueError("pattern regions have unit step")
            lo.append(a)
            hi.append(max(a, b))
        dev = torch.device("cuda", torch.cuda.current_device())
        t = torch.from_numpy(np.ascontiguousarray(mesh)).to(dev)
        def arr(v):
            return (ctypes.c_int64 * len(v))(*v)

        shape_a, lo_a, hi_a = arr(mesh.shape), arr(lo), arr(hi)
        blocks = self._blocks()
        block_a = arr(blocks) if blocks is not None else None
        _native.check(_native.lib().fw_fibrosis(
            ctypes.c_void_p(t.data_ptr()), mesh.ndim, shape_a, lo_a, hi_a,
            ctypes.cast(block_a, ctypes.c_void_p) if block_a is not None else None,
            float(self.density), ctypes.c_uint64(int(seed) & (2**64 - 1)),
            ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)), "fw_fibrosis")
        mesh[...] = t.cpu().numpy()

    def _region(self):
        raise NotImplementedError

    def _codes(self, shape):
        raise NotImplementedError


class _Diffuse(FibrosisPattern):
    def _codes(self, shape):
        # every node of the grid draws, inside the region or not
        # (diffuse_2d_pattern.py:86-87, diffuse_3d_pattern.py:88-89)
        return 1 + (np.random.random(shape) <= self.density).astype(np.int8)


class _Structural(FibrosisPattern):
    def _blocks(self):
        return tuple(int(x) for x in self._lengths())

    def _codes(self, shape):
        # blocks on a lattice anchored at the region's low corner, one draw
        # per block in C order, clipped at the region's high edges
        # (structural_2d_pattern.py:97-116, structural_3d_pattern.py:96-126)
        codes = np.ones(shape, dtype=np.int8)
        lo = [r.start for r in self._region()]
        hi = [r.stop for r in self._region()]
        step = self._lengths()
        starts = [range(a, b, s) for a, b, s in zip(lo, hi, step)]

        def corners(axis, prefix):
            if axis == len(starts):
                yield prefix
                return
            for x in starts[axis]:
                yield from corners(axis + 1, prefix + (x,))

        for c in corners(0, ()):
            if random.random() <= self.density:
                block = tuple(slice(x, x + min(s, b - x)) for x, s, b in zip(c, step, hi))
                codes[block] = 2
        return codes


class Diffuse2DPattern(_Diffuse):
    """i.i.d. fibrosis with probability `density` in [x1:x2, y1:y2]
    (cpuwave2D/fibrosis/diffuse_2d_pattern.py:6-87)."""

    def __init__(self, density, x1=None, x2=None, y1=None, y2=None):
        self.density = density
        self.x1, self.x2, self.y1, self.y2 = x1, x2, y1, y2

    def _region(self):
        return (slice(self.x1, self.x2), slice(self.y1, self.y2))


class Diffuse3DPattern(_Diffuse):
    """i.i.d. fibrosis in [x1:x2, y1:y2, z1:z2]
    (cpuwave3D/fibrosis/diffuse_3d_pattern.py:6-89).  The reference's
    `generate` takes `shape` positionally, so its `apply()` raises TypeError,
    and its mesh branch (diffuse_3d_pattern.py:83) indexes z1 and z2 as two
    separate indices, which raises IndexError; here both branches use the
    region z1:z2, as its shape branch does."""

    def __init__(self, x1, x2, y1, y2, z1, z2, density):
        self.x1, self.x2, self.y1, self.y2, self.z1, self.z2 = x1, x2, y1, y2, z1, z2
        self.density = density

    def _region(self):
        return (slice(self.x1, self.x2), slice(self.y1, self.y2), slice(self.z1, self.z2))


class Structural2DPattern(_Structural):
    """Blocks of length_i x length_j, each fibrotic with probability
    `density`, tiling [x1:x2, y1:y2] (cpuwave2D/fibrosis/structural_2d_pattern.py:7-116)."""

    def __init__(self, density, length_i, length_j, x1, x2, y1, y2):
        self.density = density
        self.length_i, self.length_j = length_i, length_j
        self.x1, self.x2, self.y1, self.y2 = x1, x2, y1, y2

    def _region(self):
        return (slice(self.x1, self.x2), slice(self.y1, self.y2))

    def _lengths(self):
        return (self.length_i, self.length_j)


class Structural3DPattern(_Structural):
    """Blocks of length_i x length_j x length_k tiling [x1:x2, y1:y2, z1:z2]
    (cpuwave3D/fibrosis/structural_3d_pattern.py:7-126)."""

    def __init__(self, x1, x2, y1, y2, z1, z2, density, length_i, length_j, length_k):
        self.x1, self.x2, self.y1, self.y2, self.z1, self.z2 = x1, x2, y1, y2, z1, z2
        self.density = density
        self.length_i, self.length_j, self.length_k = length_i, length_j, length_k

    def _region(self):
        return (slice(self.x1, self.x2), slice(self.y1, self.y2), slice(self.z1, self.z2))

    def _lengths(self):
        return (self.length_i, self.length_j, self.length_k)
