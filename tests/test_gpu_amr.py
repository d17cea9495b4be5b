"""GPU average_down (cmp_level_average_down through the C ABI) against the oracle: the coarse level
after averaging a fine level onto it is bit-identical (2D ratio 2 and 4, 3D ratio 2, two-phase and
three-phase states, fine boxes spread over several coarse boxes); cells outside the fine boxes keep
their values."""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, Level, make_geometry, make_params, split_domain

pytestmark = pytest.mark.gpu

FINE = {(2, 2): [([4, 6], [19, 13]), ([20, 2], [27, 29]), ([0, 30], [31, 31])],
        (2, 4): [([4, 8], [19, 23]), ([20, 0], [31, 15])],
        (3, 2): [([2, 4, 0], [9, 11, 7]), ([10, 0, 8], [17, 7, 15])]}


@pytest.mark.parametrize("dim,ratio,nphase", [(2, 2, 0), (2, 4, 0), (3, 2, 0), (2, 2, 3)])
def test_average_down_matches_oracle(gpu, dim, ratio, nphase):
    import torch
    nc = 16 if (dim, ratio) != (2, 4) else 8
    P = (make_params(dim, gamma=(1.4, 1.67, 1.09), nphase=3) if nphase else make_params(dim, gamma=(1.4, 1.6)))
    cg = make_geometry((nc,) * dim, (0.0,) * dim, (1.0,) * dim, (1,) * dim)
    fg = make_geometry((nc * ratio,) * dim, (0.0,) * dim, (1.0,) * dim, (1,) * dim)
    cl = Layout(dim, split_domain((nc,) * dim, max(4, nc // 2)), 3, P.nstate)
    fl = Layout(dim, FINE[(dim, ratio)], 3, P.nstate)
    rng = np.random.default_rng(7)
    fbuf = rng.uniform(0.1, 2.0, fl.total)
    cbuf = rng.uniform(-1.0, 0.0, cl.total)
    ref = cbuf.copy()
    oracle.level_average_down(P, fl, fbuf, cl, ref, ratio)
    Uf = torch.from_numpy(fbuf).to(gpu)
    Uc = torch.from_numpy(cbuf).to(gpu)
    cf, cc = Level(P, fg, fl), Level(P, cg, cl)
    cf.AverageDownTo(Uf, cc, Uc, ratio)
    assert cc.check_errors() == 0
    got = Uc.cpu().numpy()
    cf.close()
    cc.close()
    assert np.array_equal(got, ref)
    assert (got != cbuf).sum() > 0


# fine boxes (fine cells) for the two-level fill: touching domain faces, touching each other, and
# apart, so that ghosts come from fine neighbours, the coarse level and the physical BCs
FILL2 = {(2, 2): [([0, 8], [15, 23]), ([16, 8], [23, 31]), ([24, 0], [31, 7])],
         (2, 4): [([8, 0], [23, 15]), ([24, 16], [31, 31])],
         (3, 2): [([0, 4, 0], [7, 11, 7]), ([8, 4, 0], [15, 11, 7]), ([4, 12, 8], [11, 15, 15])]}
BCS = {2: dict(per=(0, 1), lo_bc=[4, 0], hi_bc=[2, 0]),
       3: dict(per=(0, 0, 1), lo_bc=[5, 3, 0], hi_bc=[4, 2, 0])}


@pytest.mark.parametrize("dim,ratio,nphase", [(2, 2, 0), (2, 4, 0), (3, 2, 0), (2, 2, 3), (3, 2, 2)])
def test_fill_two_levels_matches_oracle(gpu, dim, ratio, nphase):
    import torch
    nc = 8 if (dim, ratio) == (2, 4) or dim == 3 else 16
    if nphase:
        P = make_params(dim, gamma=(1.4, 1.67, 1.09)[:nphase], nphase=nphase)
    else:
        P = make_params(dim, gamma=(1.4, 1.6))
    bc = BCS[dim]
    cg = make_geometry((nc,) * dim, (0.0,) * dim, (1.0,) * dim, bc["per"], lo_bc=bc["lo_bc"], hi_bc=bc["hi_bc"])
    fg = make_geometry((nc * ratio,) * dim, (0.0,) * dim, (1.0,) * dim, bc["per"], lo_bc=bc["lo_bc"],
                       hi_bc=bc["hi_bc"])
    cl = Layout(dim, split_domain((nc,) * dim, max(4, nc // 2)), 3, P.nstate)
    fl = Layout(dim, FILL2[(dim, ratio)], 3, P.nstate)
    rng = np.random.default_rng(11 + dim + ratio + nphase)
    cbuf = rng.uniform(0.2, 2.0, cl.total)
    fbuf = rng.uniform(0.2, 2.0, fl.total)
    ref = fbuf.copy()
    oracle.level_fill_two_levels(P, fg, fl, ref, cg, cl, cbuf, ratio)
    Uf = torch.from_numpy(fbuf).to(gpu)
    Uc = torch.from_numpy(cbuf).to(gpu)
    lf, lc = Level(P, fg, fl), Level(P, cg, cl)
    lf.FillPatchTwoLevels(Uf, lc, Uc, ratio)
    assert lf.check_errors() == 0
    got = Uf.cpu().numpy()
    lf.close()
    lc.close()
    assert np.array_equal(got, ref)
    assert (got != fbuf).sum() > 0


@pytest.mark.parametrize("dim", [2, 3])
def test_fill_and_average_down_captured_without_tables(gpu, dim):
    """The two-level fill and the average-down keep per pair of levels a table of each cell's sources,
    built by the first call that is not under stream capture. Captured into a graph before any call
    has built them, they take their per-call source searches instead; replayed (several times, with
    eager work between the replays) they equal the table forms bit for bit."""
    import torch
    ratio = 2
    nc = 16 if dim == 2 else 8
    P = make_params(dim, gamma=(1.4, 1.6))
    bc = BCS[dim]
    cg = make_geometry((nc,) * dim, (0.0,) * dim, (1.0,) * dim, bc["per"], lo_bc=bc["lo_bc"], hi_bc=bc["hi_bc"])
    fg = make_geometry((nc * ratio,) * dim, (0.0,) * dim, (1.0,) * dim, bc["per"], lo_bc=bc["lo_bc"],
                       hi_bc=bc["hi_bc"])
    cl = Layout(dim, split_domain((nc,) * dim, max(4, nc // 2)), 3, P.nstate)
    fl = Layout(dim, FILL2[(dim, ratio)], 3, P.nstate)
    rng = np.random.default_rng(5 + dim)
    cbuf = rng.uniform(0.2, 2.0, cl.total)
    fbuf = rng.uniform(0.2, 2.0, fl.total)
    # the table forms: eager calls on one pair of levels
    lf, lc = Level(P, fg, fl), Level(P, cg, cl)
    Uf_t = torch.from_numpy(fbuf).to(gpu)
    Uc_t = torch.from_numpy(cbuf).to(gpu)
    lf.FillPatchTwoLevels(Uf_t, lc, Uc_t, ratio)
    lf.AverageDownTo(Uf_t, lc, Uc_t, ratio)
    assert lf.check_errors() == 0 and lc.check_errors() == 0
    assert not torch.equal(Uf_t, torch.from_numpy(fbuf).to(gpu))
    assert not torch.equal(Uc_t, torch.from_numpy(cbuf).to(gpu))
    # the search forms: a fresh pair of levels, captured first
    lf2, lc2 = Level(P, fg, fl), Level(P, cg, cl)
    Uf0 = torch.from_numpy(fbuf).to(gpu)
    Uc0 = torch.from_numpy(cbuf).to(gpu)
    Uf_g, Uc_g = Uf0.clone(), Uc0.clone()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        cur = torch.cuda.current_stream()
        lf2.FillPatchTwoLevels(Uf_g, lc2, Uc_g, ratio, stream=cur)
        lf2.AverageDownTo(Uf_g, lc2, Uc_g, ratio, stream=cur)
    for rep in range(3):
        Uf_g.copy_(Uf0)
        Uc_g.copy_(Uc0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(Uf_g, Uf_t), rep
        assert torch.equal(Uc_g, Uc_t), rep
    assert lf2.check_errors() == 0 and lc2.check_errors() == 0
    del g
    for lv in (lf, lc, lf2, lc2):
        lv.close()


# ---- the stage-coupled step of a hierarchy (cmp_amr_*) ---------------------------------------------

def _hierarchy(dim, nc, fine_boxes, seed=3, third=None, **kw):
    from compas_amd import cases
    h = cases.two_level_smooth(dim, nc, seed, fine_boxes, **kw)
    P = h["params"]
    geoms, boxes = list(h["geoms"]), list(h["boxes"])
    fields = list(h["fields"])
    if third is not None:
        f3 = (cases.nphase_smooth(dim, nc * 4, seed, nphase=kw["nphase"]) if kw.get("nphase")
              else cases.random_smooth(dim, nc * 4, seed))
        geoms.append(f3["geom"])
        boxes.append(third)
        fields.append(f3["U"])
    lays = [Layout(dim, b, 3, P.nstate) for b in boxes]
    bufs = [lay.pack(U) for lay, U in zip(lays, fields)]
    # the coarser levels under a finer one hold its means, as after a regrid
    for l in range(len(lays) - 2, -1, -1):
        oracle.level_average_down(P, lays[l + 1], bufs[l + 1], lays[l], bufs[l], 2)
    return P, geoms, lays, bufs


def _valid_cells(lay, buf):
    g = lay.ng
    out = []
    for b in range(lay.nbox):
        v = lay.fab_view(buf, b)
        out.append(v[(slice(None),) + (slice(g, -g),) * lay.dim].reshape(v.shape[0], -1))
    return np.concatenate(out, axis=1)


def _maxrel(lays, got, ref):
    out = []
    for lay, a, b in zip(lays, got, ref):
        x, y = _valid_cells(lay, a), _valid_cells(lay, b)
        out.append((np.abs(x - y).max(axis=1) / np.abs(y).max(axis=1)).max())
    return np.array(out)


def _assert_close(lays, got, ref, tol, run_oracle, bufs):
    """Per level, the largest per-variable max difference relative to the variable's largest
    magnitude is at most tol, or within 4x what the oracle itself moves when its input takes a
    relative perturbation of 1e-15. (The coarse-fine interpolation limits all components of a cell
    with one factor, the smallest over the components of sx/dc; at an extremum of a nearly flat
    component dc is a few ulps, its sign decides whether that factor is 0 or 1, and rounding-level
    differences of the stage state then move the fine ghost cells, and the step, by more than
    rounding.)"""
    err = _maxrel(lays, got, ref)
    if err.max() <= tol:
        return
    rng = np.random.default_rng(1)
    pert = [b * (1.0 + 1e-15 * rng.standard_normal(b.shape)) for b in bufs]
    run_oracle(pert)
    sens = _maxrel(lays, pert, ref)
    assert np.all(err <= np.maximum(tol, 4.0 * sens)), (err, sens)


AMR_CASES = {
    "2d-centre": (2, 16, [([8, 8], [23, 23])], None),
    "2d-two-boxes": (2, 16, [([4, 8], [15, 19]), ([16, 4], [27, 23])], None),
    "2d-periodic-faces": (2, 16, [([0, 8], [7, 23]), ([24, 8], [31, 15])], None),
    "2d-three-levels": (2, 16, [([8, 8], [23, 23])], [([24, 24], [39, 39])]),
    "3d-centre": (3, 8, [([4, 4, 4], [11, 11, 11])], None),
    "2d-nphase3": (2, 16, [([4, 8], [15, 19]), ([16, 4], [27, 23])], None, 3),
    "2d-nphase2-three-levels": (2, 16, [([8, 8], [23, 23])], [([24, 24], [39, 39])], 2),
}


@pytest.mark.parametrize("name", sorted(AMR_CASES))
def test_amr_step_matches_oracle(gpu, name):
    import torch
    from compas_amd import Amr
    dim, nc, fine, third, *np_ = AMR_CASES[name]
    P, geoms, lays, bufs = _hierarchy(dim, nc, fine, third=third, nphase=(np_[0] if np_ else 0))
    dt = 1.0e-3
    ref = [b.copy() for b in bufs]
    oracle.amr_rk3_step(P, geoms, lays, ref, dt, 2)
    levels = [Level(P, g, lay) for g, lay in zip(geoms, lays)]
    amr = Amr(levels, 2)
    assert amr.cf_cells(0) > 0
    Un = [torch.from_numpy(b).to(gpu) for b in bufs]
    U1 = [torch.zeros_like(u) for u in Un]
    U2 = [torch.zeros_like(u) for u in Un]
    dtd = torch.full((1,), dt, dtype=torch.float64, device=gpu)
    cmax = torch.zeros(len(levels), dtype=torch.float64, device=gpu)
    amr.rk3_step(Un, U1, U2, dtd, cmax)
    assert amr.check_errors() == 0
    got = [u.cpu().numpy() for u in Un]
    amr.close()
    for lv in levels:
        lv.close()
    _assert_close(lays, got, ref, 1e-12, lambda u: oracle.amr_rk3_step(P, geoms, lays, u, dt, 2), bufs)
    assert np.abs(_valid_cells(lays[0], got[0]) - _valid_cells(lays[0], bufs[0])).max() > 1e-8


@pytest.mark.parametrize("name,use_graph", [("2d-two-boxes", True), ("2d-three-levels", True),
                                            ("2d-three-levels", False), ("2d-nphase3", True)])
def test_amr_advance_matches_oracle_and_conserves(gpu, name, use_graph):
    import torch
    from compas_amd import Amr
    dim, nc, fine, third, *np_ = AMR_CASES[name]
    P, geoms, lays, bufs = _hierarchy(dim, nc, fine, third=third, nphase=(np_[0] if np_ else 0))
    nsteps = 20
    ref = [b.copy() for b in bufs]
    hist = oracle.amr_advance(P, geoms, lays, ref, nsteps, 0.4, 2)
    levels = [Level(P, g, lay) for g, lay in zip(geoms, lays)]
    amr = Amr(levels, 2)
    Un = [torch.from_numpy(b).to(gpu) for b in bufs]
    U1 = [torch.zeros_like(u) for u in Un]
    U2 = [torch.zeros_like(u) for u in Un]
    dtd = torch.zeros(1, dtype=torch.float64, device=gpu)
    cmax = torch.zeros(len(levels), dtype=torch.float64, device=gpu)
    dth = torch.zeros(nsteps, dtype=torch.float64, device=gpu)
    amr.advance(Un, U1, U2, nsteps, 0.4, dtd, cmax, dth, use_graph=use_graph)
    assert amr.check_errors() == 0
    got = [u.cpu().numpy() for u in Un]
    gh = dth.cpu().numpy()
    amr.close()
    for lv in levels:
        lv.close()
    assert np.all(np.abs(gh - hist) <= 1e-12 * hist), (gh, hist)
    _assert_close(lays, got, ref, 1e-10, lambda u: oracle.amr_advance(P, geoms, lays, u, nsteps, 0.4, 2), bufs)
    # composite totals (the coarse level holds the means of the finer ones) conserved
    c0, c1 = _valid_cells(lays[0], bufs[0]), _valid_cells(lays[0], got[0])
    ncons = P.nstate - (P.nphase if P.nphase else 1)     # masses, momenta, energy
    for n in range(ncons):
        assert abs(c1[n].sum() - c0[n].sum()) <= 1e-13 * np.abs(c0[n]).sum(), n


# ---- a regridded level's data (cmp_level_fill_from) --------------------------------------------------

REMAKE = {2: ([([8, 0], [23, 15]), ([24, 16], [31, 31])], [([12, 4], [27, 19]), ([0, 24], [15, 31])]),
          3: ([([0, 4, 0], [7, 11, 7]), ([8, 4, 0], [15, 11, 7])], [([4, 0, 4], [11, 7, 11]), ([4, 8, 0], [11, 15, 7])])}


@pytest.mark.parametrize("dim,nphase,with_old", [(2, 0, True), (2, 0, False), (3, 0, True), (2, 3, True)])
def test_fill_from_matches_oracle(gpu, dim, nphase, with_old):
    import torch
    nc = 8 if dim == 3 else 16
    if nphase:
        P = make_params(dim, gamma=(1.4, 1.67, 1.09)[:nphase], nphase=nphase)
    else:
        P = make_params(dim, gamma=(1.4, 1.6))
    bc = BCS[dim]
    cg = make_geometry((nc,) * dim, (0.0,) * dim, (1.0,) * dim, bc["per"], lo_bc=bc["lo_bc"], hi_bc=bc["hi_bc"])
    fg = make_geometry((nc * 2,) * dim, (0.0,) * dim, (1.0,) * dim, bc["per"], lo_bc=bc["lo_bc"], hi_bc=bc["hi_bc"])
    cl = Layout(dim, split_domain((nc,) * dim, max(4, nc // 2)), 3, P.nstate)
    new_boxes, old_boxes = REMAKE[dim]
    nl = Layout(dim, new_boxes, 3, P.nstate)
    ol = Layout(dim, old_boxes, 3, P.nstate)
    rng = np.random.default_rng(21 + dim + nphase)
    cbuf = rng.uniform(0.2, 2.0, cl.total)
    obuf = rng.uniform(0.2, 2.0, ol.total)
    ref = np.zeros(nl.total)
    oracle.level_fill_from(P, fg, nl, ref, ol if with_old else None, obuf if with_old else None, cg, cl, cbuf, 2)
    Un = torch.zeros(nl.total, dtype=torch.float64, device=gpu)
    Uo = torch.from_numpy(obuf).to(gpu)
    Uc = torch.from_numpy(cbuf).to(gpu)
    ln, lo_, lc = Level(P, fg, nl), Level(P, fg, ol), Level(P, cg, cl)
    ln.FillFrom(Un, lo_ if with_old else None, Uo if with_old else None, lc, Uc, 2)
    assert ln.check_errors() == 0
    got = Un.cpu().numpy()
    for lv in (ln, lo_, lc):
        lv.close()
    assert np.array_equal(got, ref)


def test_regrid_cycle_keeps_the_hierarchy_conservative(gpu):
    """Step a two-level hierarchy, move the fine level to new boxes (remake_level), rebuild the
    hierarchy and step on: the composite totals stay put across the regrid where the new fine
    level covers what the old one did plus coarse-only cells (its new cells interpolate the coarse
    level conservatively, its old cells keep their data)."""
    import torch
    from compas_amd import Amr, regrid
    P, geoms, lays, bufs = _hierarchy(2, 16, [([8, 8], [23, 23])])
    levels = [Level(P, g, lay) for g, lay in zip(geoms, lays)]
    Un = [torch.from_numpy(b).to(gpu) for b in bufs]
    dtd = torch.zeros(1, dtype=torch.float64, device=gpu)
    cmax = torch.zeros(2, dtype=torch.float64, device=gpu)

    def run(levels, Un, n):
        amr = Amr(levels, 2)
        S1 = [torch.zeros_like(u) for u in Un]
        S2 = [torch.zeros_like(u) for u in Un]
        amr.advance(Un, S1, S2, n, 0.4, dtd, cmax)
        assert amr.check_errors() == 0
        amr.close()

    c0 = _valid_cells(lays[0], Un[0].cpu().numpy())
    run(levels, Un, 5)
    new_fine, new_lay, Unew = regrid.remake_level(P, geoms[1], [([8, 4], [27, 23])], levels[1], Un[1], levels[0],
                                                  Un[0], 2)
    assert new_fine.check_errors() == 0
    new_fine.AverageDownTo(Unew, levels[0], Un[0], 2)
    levels[1].close()
    levels, Un = [levels[0], new_fine], [Un[0], Unew]
    run(levels, Un, 5)
    c1 = _valid_cells(lays[0], Un[0].cpu().numpy())
    for lv in levels:
        lv.close()
    for n in range(5):
        assert abs(c1[n].sum() - c0[n].sum()) <= 1e-13 * np.abs(c0[n]).sum(), n
    assert np.abs(c1 - c0).max() > 1e-6
