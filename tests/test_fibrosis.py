"""Fibrosis patterns (finitewave_amd.fibrosis).

The first four tests are the reference's own (tests/test_fibrosis_2d.py,
tests/test_fibrosis_3d.py: the same shapes, regions, densities and
tolerances).  The rest fix what those leave open: which generator is drawn,
how many draws and in which order (so a seeded mesh is the reference's), the
region and the clipping of the last blocks, and ``apply`` through the
tissue's mesh setter.  Two GPU tests step pattern-built tissues against the
oracle; the last two hold the patterns drawn on the device (fw_fibrosis) to
the reference's own statistical windows, their seed, region and blocks."""
from __future__ import annotations

import random

import numpy as np
import pytest

import finitewave_amd as fw


def _ratio(mesh, region):
    sub = mesh[region]
    return np.sum(sub == 2) / sub.size


# -- the reference's tests -----------------------------------------------------

def test_diffuse_fibrosis_2d():
    random.seed(0)
    np.random.seed(0)
    shape, (x1, x2, y1, y2) = (1000, 1000), (100, 900, 200, 800)
    res = fw.Diffuse2DPattern(density=0.3, x1=x1, x2=x2, y1=y1, y2=y2).generate(shape=shape)
    assert res.shape == shape and np.all(np.isin(res, [1, 2]))
    assert abs(_ratio(res, np.s_[x1:x2, y1:y2]) - 0.3) < 0.01


def test_structural_fibrosis_2d():
    random.seed(0)
    shape, (x1, x2, y1, y2) = (1000, 1000), (100, 900, 200, 800)
    res = fw.Structural2DPattern(density=0.4, length_i=5, length_j=4,
                                 x1=x1, x2=x2, y1=y1, y2=y2).generate(shape=shape)
    assert res.shape == shape and np.all(np.isin(res, [1, 2]))
    assert abs(_ratio(res, np.s_[x1:x2, y1:y2]) - 0.4) < 0.05


def test_diffuse_fibrosis_3d():
    random.seed(0)
    np.random.seed(0)
    shape = (100, 100, 100)
    res = fw.Diffuse3DPattern(density=0.3, x1=10, x2=90, y1=20, y2=80, z1=30,
                              z2=70).generate(shape=shape)
    assert res.shape == shape and np.all(np.isin(res, [1, 2]))
    assert abs(_ratio(res, np.s_[10:90, 20:80, 30:70]) - 0.3) < 0.01


def test_structural_fibrosis_3d():
    random.seed(0)
    shape = (100, 100, 100)
    res = fw.Structural3DPattern(density=0.4, length_i=5, length_j=4, length_k=3, x1=10,
                                 x2=90, y1=20, y2=80, z1=30, z2=70).generate(shape=shape)
    assert res.shape == shape and np.all(np.isin(res, [1, 2]))
    assert abs(_ratio(res, np.s_[10:90, 20:80, 30:70]) - 0.4) < 0.05


# -- the random streams --------------------------------------------------------

def test_diffuse_draws_the_whole_grid_from_numpy():
    """One numpy.random.random draw per node of the whole grid, in C order;
    the region then takes its slice of them.  The python generator is left
    alone.  Drawing only the region's nodes would give a different mesh and
    leave numpy's stream elsewhere."""
    shape = (13, 17)
    np.random.seed(5)
    random.seed(9)
    py_before = random.getstate()
    got = fw.Diffuse2DPattern(0.35, 2, 9, 4, 15).generate(shape=shape)
    after = np.random.random()
    assert random.getstate() == py_before
    np.random.seed(5)
    full = np.random.random(shape)
    assert np.random.random() == after
    want = np.ones(shape, dtype=np.int8)
    want[2:9, 4:15] = np.where(full[2:9, 4:15] <= 0.35, 2, 1)
    assert got.dtype == np.int8 and np.array_equal(got, want)


@pytest.mark.parametrize("dim", [2, 3])
def test_structural_draws_one_python_random_per_block_in_c_order(dim):
    """Block anchors run from the region's low corner in steps of the block
    lengths, the first axis slowest; block b is fibrosis iff the b-th
    random.random() is <= density.  Checked node by node with anchors built
    by numpy's meshgrid, so a transposed loop order or a wrong anchor fails."""
    if dim == 2:
        shape, lo, hi, ln = (23, 29), (3, 5), (21, 26), (4, 3)
        pat = fw.Structural2DPattern(0.5, ln[0], ln[1], lo[0], hi[0], lo[1], hi[1])
    else:
        shape, lo, hi, ln = (14, 12, 13), (1, 2, 3), (13, 11, 12), (3, 4, 2)
        pat = fw.Structural3DPattern(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 0.5,
                                     ln[0], ln[1], ln[2])
    counts = [-(-(h - l) // s) for l, h, s in zip(lo, hi, ln)]
    n_blocks = int(np.prod(counts))
    random.seed(11)
    draws = np.array([random.random() for _ in range(n_blocks + 1)])
    random.seed(11)
    np.random.seed(0)
    np_before = np.random.get_state()[1].copy()
    got = pat.generate(shape=shape)
    assert random.random() == draws[n_blocks]           # exactly n_blocks draws
    assert np.array_equal(np.random.get_state()[1], np_before)
    fib = (draws[:n_blocks] <= 0.5).reshape(counts)     # C order: first axis slowest
    assert 0 < fib.sum() < n_blocks
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    inside = np.ones(shape, dtype=bool)
    block = []
    for a in range(dim):
        inside &= (idx[a] >= lo[a]) & (idx[a] < hi[a])
        block.append(np.clip((idx[a] - lo[a]) // ln[a], 0, counts[a] - 1))
    want = np.where(inside & fib[tuple(block)], 2, 1)
    assert np.array_equal(got, want)


def test_structural_blocks_are_clipped_at_the_region_edge():
    """Density 1 fills the region exactly: no block runs past x2 or y2 even
    when the lengths do not divide the region."""
    random.seed(0)
    got = fw.Structural2DPattern(1.0, 7, 5, 2, 17, 3, 14).generate(shape=(20, 20))
    want = np.ones((20, 20), dtype=np.int8)
    want[2:17, 3:14] = 2
    assert np.array_equal(got, want)
    got3 = fw.Structural3DPattern(1, 8, 2, 9, 0, 5, 1.0, 4, 4, 4).generate(shape=(10, 10, 6))
    want3 = np.ones((10, 10, 6), dtype=np.int8)
    want3[1:8, 2:9, 0:5] = 2
    assert np.array_equal(got3, want3)


# -- generate(mesh=...) and apply ----------------------------------------------

def test_generate_needs_shape_or_mesh():
    with pytest.raises(ValueError):
        fw.Diffuse2DPattern(0.2).generate()
    with pytest.raises(NotImplementedError):
        fw.FibrosisPattern().generate(shape=(4, 4))


def test_generate_on_a_mesh_writes_only_the_region():
    mesh = np.zeros((12, 12), dtype=np.int8)
    np.random.seed(1)
    out = fw.Diffuse2DPattern(0.5, 3, 8, 2, 10).generate(mesh=mesh)
    assert out is mesh
    outside = np.ones_like(mesh, dtype=bool)
    outside[3:8, 2:10] = False
    assert np.all(mesh[outside] == 0) and np.all(np.isin(mesh[3:8, 2:10], [1, 2]))
    mesh3 = np.zeros((8, 9, 10), dtype=np.int8)
    fw.Diffuse3DPattern(1, 7, 2, 8, 3, 9, 0.5).generate(mesh=mesh3)
    outside3 = np.ones_like(mesh3, dtype=bool)
    outside3[1:7, 2:8, 3:9] = False
    assert np.all(mesh3[outside3] == 0) and np.all(np.isin(mesh3[1:7, 2:8, 3:9], [1, 2]))


@pytest.mark.parametrize("dim", [2, 3])
def test_add_pattern_sets_the_mesh_and_keeps_the_zero_boundary(dim):
    """tissue.add_pattern(p) = p.apply(tissue): the tissue's mesh setter
    receives the patterned mesh and puts the zero outer layer back, as the
    reference's setter does (core/tissue/cardiac_tissue.py:38-52)."""
    shape = [16, 18] if dim == 2 else [9, 10, 11]
    tissue = (fw.CardiacTissue2D if dim == 2 else fw.CardiacTissue3D)(shape)
    np.random.seed(2)
    pat = fw.Diffuse2DPattern(0.3) if dim == 2 else \
        fw.Diffuse3DPattern(0, shape[0], 0, shape[1], 0, shape[2], 0.3)
    tissue.add_pattern(pat)
    m = tissue.mesh
    inner = (slice(1, -1),) * dim
    assert np.all(np.isin(m[inner], [1, 2])) and np.any(m[inner] == 2)
    boundary = np.ones(m.shape, dtype=bool)
    boundary[inner] = False
    assert np.all(m[boundary] == 0)
    tissue.check_boundary()
    tissue.clean()
    assert not np.any(tissue.mesh == 2)


@pytest.mark.gpu
def test_interstitial_fibrosis_run_matches_oracle(gpu):
    """The interstitial-fibrosis setup (Structural2DPattern, 1x4 blocks at
    20 %): a planar wave through it, 1,000 Aliev-Panfilov steps, bit for bit
    the oracle's on the same mesh."""
    from oracle.oracle import OracleSim, OStim
    n = 80
    tissue = fw.CardiacTissue2D([n, n])
    random.seed(4)
    tissue.add_pattern(fw.Structural2DPattern(density=0.2, length_i=1, length_j=4,
                                              x1=0, x2=n, y1=0, y2=n))
    assert 0.1 < np.mean(tissue.mesh[1:-1, 1:-1] == 2) < 0.3
    model = fw.AlievPanfilov2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 10
    model.cardiac_tissue = tissue
    model.prog_bar = False
    seq = fw.StimSequence()
    seq.add_stim(fw.StimVoltageCoord2D(0, 1.0, 0, n, 0, 5))
    model.stim_sequence = seq
    model.run()
    sim = OracleSim("aliev_panfilov", tissue.mesh, 0.01, 0.25)
    sim.stims.append(OStim("voltage", 0, 1.0, (slice(0, n), slice(0, 5))))
    sim.run(10)
    assert np.array_equal(model.u, sim.u)
    assert np.array_equal(model.v, sim.state["v"])


@pytest.mark.gpu
def test_fibrosis_added_by_a_command_mid_run_matches_oracle(gpu):
    """A command at t = 4 writes a structural pattern into the mesh and calls
    compute_weights(); the wave then meets the new obstacle.  Bit for bit
    the oracle's run with the same command, and a second run() continues on
    the new mesh."""
    from oracle.oracle import OCommand, OracleSim, OStim
    n = 60
    random.seed(7)
    block = fw.Structural2DPattern(0.5, 3, 3, 20, 44, 10, 50).generate(shape=(n, n))

    class AddFibrosis(fw.Command):
        def execute(self, model):
            model.cardiac_tissue.mesh = np.where(block == 2, 2, model.cardiac_tissue.mesh
                                                 ).astype(np.int8)
            model.compute_weights()

    tissue = fw.CardiacTissue2D([n, n])
    model = fw.AlievPanfilov2D()
    model.dt, model.dr, model.t_max = 0.01, 0.25, 10
    model.cardiac_tissue = tissue
    model.prog_bar = False
    seq = fw.StimSequence()
    seq.add_stim(fw.StimVoltageCoord2D(0, 1.0, 0, n, 0, 4))
    model.stim_sequence = seq
    cs = fw.CommandSequence()
    cs.add_command(AddFibrosis(4))
    model.command_sequence = cs
    model.run()
    assert np.any(model.cardiac_tissue.mesh == 2)

    mesh0 = fw.CardiacTissue2D([n, n]).mesh

    def add(sim):
        sim.mesh = np.where(block == 2, 2, sim.mesh).astype(np.int8)
        sim.mesh[0, :] = sim.mesh[-1, :] = sim.mesh[:, 0] = sim.mesh[:, -1] = 0
        sim.compute_weights()

    sim = OracleSim("aliev_panfilov", mesh0, 0.01, 0.25)
    sim.stims.append(OStim("voltage", 0, 1.0, (slice(0, n), slice(0, 4))))
    sim.commands.append(OCommand(4, add))
    sim.run(10)
    assert np.array_equal(sim.mesh, model.cardiac_tissue.mesh)
    assert np.array_equal(model.u, sim.u)
    assert np.array_equal(model.v, sim.state["v"])
    model.t_max = 14
    model.run(initialize=False)
    sim.run(14, initialize=False)
    assert np.array_equal(model.u, sim.u)


# -- generation on the device --------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("case", ["diffuse2d", "structural2d", "diffuse3d", "structural3d"])
def test_device_patterns_meet_the_reference_windows(gpu, case):
    """The reference's four fibrosis tests (same shapes, regions, densities,
    tolerances) with the pattern drawn on the device."""
    if case == "diffuse2d":
        pat, shape, region, dens, tol = (fw.Diffuse2DPattern(0.3, 100, 900, 200, 800),
                                         (1000, 1000), np.s_[100:900, 200:800], 0.3, 0.01)
    elif case == "structural2d":
        pat, shape, region, dens, tol = (fw.Structural2DPattern(0.4, 5, 4, 100, 900, 200, 800),
                                         (1000, 1000), np.s_[100:900, 200:800], 0.4, 0.05)
    elif case == "diffuse3d":
        pat, shape, region, dens, tol = (fw.Diffuse3DPattern(10, 90, 20, 80, 30, 70, 0.3),
                                         (100, 100, 100), np.s_[10:90, 20:80, 30:70], 0.3, 0.01)
    else:
        pat, shape, region, dens, tol = (fw.Structural3DPattern(10, 90, 20, 80, 30, 70, 0.4,
                                                                5, 4, 3),
                                         (100, 100, 100), np.s_[10:90, 20:80, 30:70], 0.4, 0.05)
    res = pat.generate(shape=shape, device=True, seed=0)
    assert res.shape == shape and res.dtype == np.int8 and np.all(np.isin(res, [1, 2]))
    assert abs(_ratio(res, region) - dens) < tol
    outside = np.ones(shape, dtype=bool)
    outside[region] = False
    assert np.all(res[outside] == 1)


@pytest.mark.gpu
def test_device_patterns_are_keyed_by_seed_and_keep_blocks_whole(gpu):
    """The same seed gives the same mesh, another seed another; the nodes
    outside the region keep their codes; every structural block is all
    tissue or all fibrosis, clipped at the region's high edges."""
    pat = fw.Diffuse2DPattern(0.5, 3, 40, 5, 33)
    a = pat.generate(shape=(44, 37), device=True, seed=7)
    b = pat.generate(shape=(44, 37), device=True, seed=7)
    c = pat.generate(shape=(44, 37), device=True, seed=8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    mesh = np.zeros((44, 37), dtype=np.int8)
    out = pat.generate(mesh=mesh, device=True, seed=7)
    assert out is mesh
    assert np.array_equal(mesh[3:40, 5:33], a[3:40, 5:33])
    rest = np.ones(mesh.shape, dtype=bool)
    rest[3:40, 5:33] = False
    assert np.all(mesh[rest] == 0)
    # structural 3D: blocks 4 x 3 x 5 from (2, 1, 0), region ends cut the last ones
    lo, hi, ln = (2, 1, 0), (19, 15, 13), (4, 3, 5)
    s = fw.Structural3DPattern(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 0.5, *ln)
    m = s.generate(shape=(20, 16, 14), device=True, seed=3)
    n_blocks, n_fib = 0, 0
    for bi in range(lo[0], hi[0], ln[0]):
        for bj in range(lo[1], hi[1], ln[1]):
            for bk in range(lo[2], hi[2], ln[2]):
                blk = m[bi:min(bi + ln[0], hi[0]), bj:min(bj + ln[1], hi[1]),
                        bk:min(bk + ln[2], hi[2])]
                assert np.all(blk == blk.flat[0])
                n_blocks += 1
                n_fib += int(blk.flat[0] == 2)
    assert 0 < n_fib < n_blocks
    # a tissue takes the device pattern through apply(), keeping its zero boundary
    tissue = fw.CardiacTissue2D([30, 30])
    fw.Diffuse2DPattern(0.3).apply(tissue, device=True, seed=1)
    assert np.all(tissue.mesh[0] == 0) and np.any(tissue.mesh[1:-1, 1:-1] == 2)


def test_device_entry_point_checks_arguments_before_any_launch():
    """fw_fibrosis refuses a bad rank, a region outside the grid or inverted,
    and a block length below 1, with a negative status and a message naming
    the axis; an empty region returns 0.  These paths return before any
    device call, so they run without a GPU (the mesh pointer is never read)."""
    import ctypes

    from finitewave_amd import _native
    L = _native.lib()

    def arr(v):
        return (ctypes.c_int64 * len(v))(*v)

    mesh = ctypes.c_void_p(256)

    def call(ndim, shape, lo, hi, block=None):
        blk = ctypes.cast(arr(block), ctypes.c_void_p) if block is not None else None
        return L.fw_fibrosis(mesh, ndim, arr(shape), arr(lo), arr(hi), blk, 0.3,
                             ctypes.c_uint64(5), None)

    assert call(4, [5, 5, 5], [0, 0, 0], [5, 5, 5]) < 0
    assert b"bad arguments" in L.fw_last_error()
    assert call(2, [5, 5], [0, 1], [5, 6]) < 0          # past the grid's end
    assert b"axis 1" in L.fw_last_error()
    assert call(2, [5, 5], [3, 0], [2, 5]) < 0          # lo > hi
    assert b"axis 0" in L.fw_last_error()
    assert call(2, [5, 5], [-1, 0], [2, 5]) < 0         # negative start
    assert b"axis 0" in L.fw_last_error()
    assert call(3, [5, 5, 5], [0, 0, 0], [5, 5, 5], [1, 0, 1]) < 0
    assert b"axis 1" in L.fw_last_error() and b"block 0" in L.fw_last_error()
    assert call(2, [5, 5], [2, 0], [2, 5]) == 0          # empty region: nothing to draw
    assert call(3, [5, 5, 5], [0, 0, 4], [5, 5, 4], [2, 2, 2]) == 0
