"""The persistent work-list stage kernel (cmp_level_set_tiles, cmp_amr_tiles) and the tile autotuner
(cmp_level_tune_tiles) through the C ABI.

Every tile shape computes each face once with the same arithmetic and each cell's update from the
same face values, so a level stepped through the work-list kernel must equal the fixed-shape kernel
bit for bit, at every shape (including shapes larger than some boxes, shapes that need the second
face pass, odd shapes) and both occupancy targets; several replayed steps also check that the
kernel's device counters are left at zero by each launch (a launch that found them non-zero would
skip tiles). One case is also checked against the oracle directly. The hierarchy's single launch
over every level's tiles must equal the per-level launches bit for bit."""
import numpy as np
import pytest

import oracle
from compas_amd import Amr, Layout, Level, cases, tile_cache_clear, tile_cache_load, tile_cache_save

pytestmark = pytest.mark.gpu


def _steps(gpu, c, nsteps, ext=None, waves=2, boxes=None):
    """nsteps graph-replayed TVD-RK3 steps of case c (its boxes or `boxes`), fixed kernel when ext is
    None; returns (state, dt history, tile shapes)."""
    import torch
    P = c["params"]
    lay = Layout(P.dim, boxes if boxes is not None else c["boxes"], 3, P.nstate)
    Un = torch.from_numpy(lay.pack(c["U"])).to(gpu)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dt = torch.zeros(1, dtype=torch.float64, device=gpu)
    cmax = torch.zeros(1, dtype=torch.float64, device=gpu)
    hist = torch.zeros(nsteps, dtype=torch.float64, device=gpu)
    lev = Level(P, c["geom"], lay)
    if ext is not None:
        lev.set_tiles(ext, waves)
    shapes = lev.tile_shapes()
    lev.advance(Un, U1, U2, nsteps, c["cfl"], dt, cmax, dt_hist=hist, use_graph=True)
    assert lev.check_errors() == 0
    out = Un.cpu().numpy()
    h = hist.cpu().numpy()
    lev.close()
    return out, h, shapes


def _check_shapes(gpu, c, shapes, boxes=None, nsteps=3):
    ref, href, sh = _steps(gpu, c, nsteps, boxes=boxes)
    assert not sh[0]
    for ext in shapes:
        for waves in (2, 3):
            got, h, sh = _steps(gpu, c, nsteps, ext=ext, waves=waves, boxes=boxes)
            assert sh[0] and sh[1] == waves
            print("shape", ext, "waves", waves, flush=True)
            assert np.array_equal(h, href), (ext, waves)
            assert np.array_equal(got, ref), (ext, waves, np.max(np.abs(got - ref)))


def test_list_kernel_two_phase_thinc_2d(gpu):
    """Two-phase WENO5 + THINC + HLLC on uneven boxes of 8..64 cells (a regrid's box list)."""
    c = cases.random_smooth(2, 128, seed=3, thinc=1)
    boxes = cases.uneven_boxes((128, 128), seed=5)
    _check_shapes(gpu, c, [(16, 15, 1), (16, 16, 1), (32, 8, 1), (8, 32, 1), (24, 10, 1), (8, 8, 1), (5, 7, 1),
                           (32, 4, 1)], boxes=boxes)


def test_list_kernel_nphase_2d(gpu):
    """Three-phase WENO5 + HLLC, the AMR configuration's base state and boxes."""
    c = cases.nphase_rings_2d(128)
    _check_shapes(gpu, c, [(16, 16, 1), (32, 8, 1), (24, 8, 1), (13, 9, 1)])


def test_list_kernel_3d(gpu):
    """3D two-phase WENO5 + THINC + HLLC, boxes of 16 cells, shapes with one and two face passes."""
    c = cases.random_smooth(3, 32, seed=11, thinc=1, max_grid_size=16)
    _check_shapes(gpu, c, [(8, 8, 3), (8, 8, 4), (16, 4, 4), (4, 8, 8), (4, 4, 16), (3, 5, 7)], nsteps=2)


@pytest.mark.parametrize("scheme,riemann", [("WENO5B", "HLLC"), ("MUSCL-MC", "HLL"), ("GODUNOV", "HLLC"),
                                            ("WENO5", "HLL")])
def test_list_kernel_schemes(gpu, scheme, riemann):
    c = cases.random_smooth(2, 64, seed=2, scheme=scheme, riemann=riemann)
    _check_shapes(gpu, c, [(16, 16, 1), (8, 24, 1)], boxes=cases.uneven_boxes((64, 64), seed=1), nsteps=2)


def test_list_kernel_matches_oracle(gpu):
    """One step through the work-list kernel against the oracle's step (1e-12 per variable)."""
    c = cases.advected_interface_2d(n=64)
    boxes = cases.uneven_boxes((64, 64), seed=4)
    got, _, _ = _steps(gpu, c, 1, ext=(16, 16, 1), waves=3, boxes=boxes)
    P = c["params"]
    lay = Layout(2, boxes, 3, P.nstate)
    buf = lay.pack(c["U"])
    oracle.level_advance(P, c["geom"], lay, buf, 1, c["cfl"])
    ref = lay.unpack(buf, c["shape"])
    g = lay.unpack(got, c["shape"])
    for n in range(P.nstate):
        assert np.max(np.abs(g[n] - ref[n])) <= 1e-12 * np.max(np.abs(ref[n])), n


def test_tune_tiles_and_cache(gpu, tmp_path):
    """The measured sweep: every box shape gets a cached entry with positive times, the level's chosen
    shapes fit its boxes, the step after tuning equals the fixed kernel's, and the cache survives a
    save / clear / load round trip."""
    import torch
    c = cases.nphase_rings_2d(128)
    P = c["params"]
    ref, href, _ = _steps(gpu, c, 3)
    lay = Layout(2, c["boxes"], 3, P.nstate)
    Un = torch.from_numpy(lay.pack(c["U"])).to(gpu)
    lev = Level(P, c["geom"], lay)
    tile_cache_clear()
    lev.tune_tiles(Un, extra_shapes=[(40, 8)], force=1)
    on, waves, ext = lev.tile_shapes()
    assert on and waves in (2, 3)
    shapes = set()
    for (lo, hi), e in zip(lay.boxes, ext):
        n = (hi[0] - lo[0] + 1, hi[1] - lo[1] + 1)
        shapes.add(n)
        assert 1 <= e[0] <= n[0] and 1 <= e[1] <= n[1] and e[0] * e[1] <= 256
    for n in list(shapes) + [(40, 8)]:
        t = lev.tile_tuning(n)
        assert t is not None and t["fixed"] > 0 and t["cand"], n
        assert all(a > 0 and b > 0 for _, a, b in t["cand"])
    path = tmp_path / "tiles.txt"
    tile_cache_save(path)
    before = {n: lev.tile_tuning(n) for n in shapes}
    tile_cache_clear()
    assert lev.tile_tuning(next(iter(shapes))) is None
    tile_cache_load(path)
    for n in shapes:
        a, b = before[n], lev.tile_tuning(n)
        assert [e for e, _, _ in a["cand"]] == [e for e, _, _ in b["cand"]]
        assert np.allclose([x for _, x, _ in a["cand"]], [x for _, x, _ in b["cand"]], rtol=1e-8)
    lev.close()
    got, h, sh = _steps(gpu, c, 3, ext=[v for e in ext for v in e], waves=waves)
    assert np.array_equal(h, href) and np.array_equal(got, ref)


def test_mixed_level_equals_fixed(gpu):
    """A level that keeps some boxes with the fixed-shape kernel (extents 0 0 0) and steps the others
    through the work-list kernel equals the fixed kernel bit for bit, and reports its choice."""
    c = cases.nphase_rings_2d(128)
    boxes = cases.uneven_boxes((128, 128), seed=7)
    ref, href, _ = _steps(gpu, c, 3, boxes=boxes)
    for waves in (2, 3):
        ext = [v for b in range(len(boxes)) for v in ((0, 0, 0) if b % 3 == 0 else (24, 8, 1))]
        got, h, (on, w, shapes) = _steps(gpu, c, 3, ext=ext, waves=waves, boxes=boxes)
        assert on and w == waves
        assert [s[0] == 0 for s in shapes] == [b % 3 == 0 for b in range(len(boxes))]
        assert np.array_equal(h, href) and np.array_equal(got, ref), waves
    # every box 0 0 0: the fixed kernel alone
    _, _, (on, _, _) = _steps(gpu, c, 1, ext=[0, 0, 0] * len(boxes), boxes=boxes)
    assert not on


def _hierarchy_steps(gpu, c, force, nsteps=3, level_tiles=None):
    """nsteps of the hierarchy; level_tiles: per level a function of the box index giving its tile
    extents (0 0 0: the fixed kernel) and the levels' own choice instead of the tuner's."""
    import torch
    P = c["params"]
    levels, Us = [], []
    for g, boxes, f in zip(c["geoms"], c["boxes"], c["fields"]):
        lay = Layout(2, boxes, 3, P.nstate)
        levels.append(Level(P, g, lay))
        Us.append(torch.from_numpy(lay.pack(f)).to(gpu))
    for l, (lv, U) in enumerate(zip(levels, Us)):
        if level_tiles is None:
            lv.tune_tiles(U)
        elif level_tiles[l] is not None:
            lv.set_tiles([v for b in range(lv.layout.nbox) for v in level_tiles[l](b)], 3)
    amr = Amr(levels, 2)
    on, waves = amr.tiles(force)
    assert on == (force > 0)
    S1 = [torch.zeros_like(u) for u in Us]
    S2 = [torch.zeros_like(u) for u in Us]
    dt = torch.zeros(1, dtype=torch.float64, device=gpu)
    cmax = torch.zeros(len(levels), dtype=torch.float64, device=gpu)
    hist = torch.zeros(nsteps, dtype=torch.float64, device=gpu)
    amr.advance(Us, S1, S2, nsteps, 0.4, dt, cmax, dt_hist=hist)
    assert amr.check_errors() == 0
    out = [u.cpu().numpy() for u in Us]
    h = hist.cpu().numpy()
    amr.close()
    for lv in levels:
        lv.close()
    return out, h


@pytest.mark.parametrize("nphase", [0, 3])
def test_hierarchy_single_list_launch_equals_per_level(gpu, nphase):
    """Three levels (ratio 2) stepped with one work-list launch per stage over every level's tuned
    tiles equal the per-level fixed-kernel launches bit for bit."""
    def blocks(lo, hi, b):
        return [([i, j], [i + b - 1, j + b - 1]) for j in range(lo, hi, b) for i in range(lo, hi, b)]

    c = cases.interface_hierarchy(2, 32, [blocks(8, 56, 16), blocks(32, 80, 16) + blocks(80, 96, 8)[:1]],
                                  nphase=nphase, rings=bool(nphase), max_grid_size=16)
    ref, href = _hierarchy_steps(gpu, c, -1)
    got, h = _hierarchy_steps(gpu, c, 1)
    assert np.array_equal(h, href)
    for l, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), (l, np.max(np.abs(a - b)))


@pytest.mark.parametrize("nphase", [0, 3])
def test_hierarchy_levels_mixed_tiles_equal_fixed(gpu, nphase):
    """Each level of a hierarchy stepped with its own tiles (level 0 fixed, level 1 mixed, level 2
    work-list only) through the per-level stages, which read the filled ghost cells, equals every
    level on the fixed kernel bit for bit."""
    def blocks(lo, hi, b):
        return [([i, j], [i + b - 1, j + b - 1]) for j in range(lo, hi, b) for i in range(lo, hi, b)]

    c = cases.interface_hierarchy(2, 32, [blocks(8, 56, 16), blocks(32, 80, 16) + blocks(80, 96, 8)[:1]],
                                  nphase=nphase, rings=bool(nphase), max_grid_size=16)
    none = [None, None, None]
    ref, href = _hierarchy_steps(gpu, c, -1, level_tiles=none)
    mixed = [None, lambda b: (0, 0, 0) if b % 2 else (8, 16, 1), lambda b: (16, 8, 1)]
    got, h = _hierarchy_steps(gpu, c, -1, level_tiles=mixed)
    assert np.array_equal(h, href)
    for l, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), (l, np.max(np.abs(a - b)))
