"""The host clustering (CPU only): the boxes cover every tag grown by n_error_buf, do not overlap,
sit on blocking-factor lines, are at most max_grid_size a side, and each box's blocks are tagged
at least grid_eff of the time or the box could not be cut further; the GPU tags of a ring feed it
the same way (cmp_level_tag's layout through tags_to_global)."""
import numpy as np
import pytest

from compas_amd import regrid


def _check(tags, boxes, bf, mgs, nbuf):
    dim = tags.ndim
    cover = np.zeros(tags.shape, dtype=int)
    for lo, hi in boxes:
        for d in range(dim):
            assert lo[d] % bf == 0 and (hi[d] + 1) % bf == 0 or hi[d] == tags.shape[dim - 1 - d] - 1
            assert hi[d] - lo[d] + 1 <= mgs
        sl = tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(dim)))
        cover[sl] += 1
    assert cover.max() <= 1
    grown = regrid._grow(tags, nbuf)
    assert np.all(cover[grown] == 1)


@pytest.mark.parametrize("dim,seed", [(2, 0), (2, 1), (3, 2)])
def test_cluster_covers_tags(dim, seed):
    rng = np.random.default_rng(seed)
    n = 128 if dim == 2 else 48
    X = np.indices((n,) * dim).astype(float) / n
    tags = np.zeros((n,) * dim, dtype=np.uint8)
    for _ in range(4):
        c = rng.uniform(0.2, 0.8, dim)
        r = np.sqrt(sum((X[d] - c[d]) ** 2 for d in range(dim)))
        tags |= (np.abs(r - rng.uniform(0.05, 0.2)) < 1.5 / n).astype(np.uint8)
    boxes = regrid.cluster(tags, blocking_factor=8, max_grid_size=32, grid_eff=0.7, n_error_buf=1)
    _check(tags, boxes, 8, 32, 1)
    cells = sum(np.prod([h - l + 1 for l, h in zip(lo, hi)]) for lo, hi in boxes)
    assert cells < tags.size                     # the clustering refines a part of the level only


def test_empty_and_full():
    assert regrid.cluster(np.zeros((32, 32), dtype=np.uint8)) == []
    full = regrid.cluster(np.ones((32, 32), dtype=np.uint8), blocking_factor=8, max_grid_size=16)
    assert sorted(full) == sorted([([i, j], [i + 15, j + 15]) for i in (0, 16) for j in (0, 16)])


def test_refine_and_tags_to_global():
    assert regrid.refine([([8, 0], [15, 7])], 2) == [([16, 0], [31, 15])]
    from compas_amd import Layout, split_domain
    lay = Layout(2, split_domain((16, 16), 8), 3, 6)
    g = np.arange(256, dtype=np.uint8).reshape(16, 16) % 2
    flat = np.concatenate([g[lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].ravel() for lo, hi in lay.boxes])
    assert np.array_equal(regrid.tags_to_global(lay, flat, (16, 16)), g)


def test_nest_keeps_whole_blocks_inside_the_parent_with_a_margin():
    parent = [([32, 32], [95, 95]), ([96, 40], [127, 79])]
    boxes = [([24, 24], [103, 55]), ([40, 64], [63, 119]), ([96, 48], [119, 71])]
    out = regrid.nest(boxes, parent, (160, 160), 8, 4)
    cov = np.zeros((160, 160), bool)
    for lo, hi in parent:
        cov[lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    want = np.zeros((160, 160), bool)
    for lo, hi in boxes:
        want[lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    m = np.zeros((160, 160), int)
    for lo, hi in out:
        assert lo[0] % 8 == 0 and lo[1] % 8 == 0 and (hi[0] + 1) % 8 == 0 and (hi[1] + 1) % 8 == 0
        assert cov[lo[1] - 4:hi[1] + 5, lo[0] - 4:hi[0] + 5].all(), (lo, hi)
        m[lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert m.max() == 1
    # nothing requested is dropped that a whole inner block could hold
    inner = np.zeros((160, 160), bool)
    for lo, hi in parent:
        inner[lo[1] + 4:hi[1] - 3, lo[0] + 4:hi[0] - 3] = True
    blk = (inner & want).reshape(20, 8, 20, 8).all(axis=(1, 3))
    got = (m > 0).reshape(20, 8, 20, 8).all(axis=(1, 3))
    assert np.all(got[blk])
    assert np.all(~(m > 0) | want)


def test_nest_in_3d_keeps_whole_blocks_inside_the_parent_with_a_margin():
    """3D nesting: boxes of whole 4-cell blocks, every box 4 cells inside the parent union, no
    overlap, at most 64 cells a side, and every requested block that lies inside kept."""
    parent = [([8, 8, 8], [39, 39, 31]), ([40, 8, 8], [55, 23, 31])]
    boxes = [([4, 4, 4], [47, 19, 27]), ([12, 20, 12], [27, 43, 23])]
    n = 64
    out = regrid.nest(boxes, parent, (n, n, n), 4, 4)
    cov = np.zeros((n, n, n), bool)
    for lo, hi in parent:
        cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    want = np.zeros((n, n, n), bool)
    for lo, hi in boxes:
        want[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    m = np.zeros((n, n, n), int)
    assert out
    for lo, hi in out:
        assert all(lo[d] % 4 == 0 and (hi[d] + 1) % 4 == 0 and hi[d] - lo[d] + 1 <= 64 for d in range(3))
        assert cov[lo[2] - 4:hi[2] + 5, lo[1] - 4:hi[1] + 5, lo[0] - 4:hi[0] + 5].all(), (lo, hi)
        m[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert m.max() == 1
    inner = ~regrid._grow(~cov, 4)
    blk = (inner & want).reshape(16, 4, 16, 4, 16, 4).all(axis=(1, 3, 5))
    got = (m > 0).reshape(16, 4, 16, 4, 16, 4).all(axis=(1, 3, 5))
    assert np.all(got[blk])
    assert np.all(~(m > 0) | want)


@pytest.mark.parametrize("dim,seed", [(2, 0), (2, 1), (2, 2), (3, 3), (3, 4)])
def test_native_split_equals_python_split(dim, seed):
    """The library's split step (cmp_cluster_blocks) gives the Python _split's boxes in its order, on
    random sparse tags, on clustered tags with holes, and with grid_eff 0 and 1."""
    from compas_amd import regrid
    rng = np.random.default_rng(seed)
    n = 64 if dim == 2 else 24
    shape = (n,) * dim
    sparse = rng.random(shape) < 0.02
    grid = np.indices(shape)
    c = np.array(shape)[:, None] / 2.0
    r = np.sqrt(((grid.reshape(dim, -1) - c) ** 2).sum(axis=0)).reshape(shape)
    rings = (np.abs(r - n * 0.2) < 1.5) | (np.abs(r - n * 0.35) < 1.0)
    for tags in (sparse, rings, sparse | rings):
        for eff in (0.0, 0.5, 0.7, 1.0):
            for bf, nbuf in ((1, 0), (2, 1), (4, 1)):
                a = regrid.cluster(tags.astype(np.uint8), blocking_factor=bf, max_grid_size=16 * bf,
                                   grid_eff=eff, n_error_buf=nbuf, native=True)
                b = regrid.cluster(tags.astype(np.uint8), blocking_factor=bf, max_grid_size=16 * bf,
                                   grid_eff=eff, n_error_buf=nbuf, native=False)
                assert a == b, (eff, bf, nbuf)
    assert len(a) > 1


def test_grow_is_the_dilation_without_wrap():
    """_grow(tags, n): a cell is tagged when a tagged cell lies within n cells along one axis, axis
    by axis (a box dilation of radius n), nothing carried across the array's ends."""
    from compas_amd import regrid
    rng = np.random.default_rng(0)
    for shape in [(7, 9), (1, 5), (2, 2), (5, 6, 7), (40, 33)]:
        t = rng.random(shape) < 0.1
        for n in (0, 1, 2, 3, 8):
            want = np.zeros(shape, dtype=bool)
            for idx in zip(*np.nonzero(t)):
                sl = tuple(slice(max(0, i - n), i + n + 1) for i in idx)
                want[sl] = True
            assert np.array_equal(regrid._grow(t, n), want), (shape, n)
