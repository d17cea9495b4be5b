"""Pins of the oracle's average_down (AverageDownTo; CPU only):

- a constant fine level gives the constant; a linear field (cell averages = centre values) gives the
  coarse centre values; the sum of fine cell volumes times values equals the coarse one under each
  fine box (conservation);
- coarse cells outside the fine boxes are untouched; a fine box not aligned to the ratio is refused.
"""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, make_params, split_domain


def _setup(dim, ratio, nc=16, fine_boxes=None, seed=0):
    P = make_params(dim, gamma=(1.4, 1.6))
    cl = Layout(dim, split_domain((nc,) * dim, nc // 2), 3, P.nstate)
    fb = fine_boxes or ([([4, 6], [19, 13]), ([20, 2], [27, 29])] if dim == 2 else
                        [([2, 4, 0], [9, 11, 7]), ([10, 0, 8], [17, 7, 15])])
    fl = Layout(dim, fb, 3, P.nstate)
    return P, cl, fl


def _fine_field(fl, dim, f):
    buf = np.zeros(fl.total)
    for b, (lo, hi) in enumerate(fl.boxes):
        v = fl.fab_view(buf, b)
        ng = fl.ng
        idx = np.indices(v.shape[1:])
        X = [idx[dim - 1 - d] + lo[d] - ng for d in range(dim)]   # cell indices per direction
        for n in range(v.shape[0]):
            v[n] = f(n, X)
    return buf


@pytest.mark.parametrize("dim,ratio", [(2, 2), (2, 4), (3, 2)])
def test_constant_linear_and_conservation(dim, ratio):
    P, cl, fl = _setup(dim, ratio)
    if ratio == 4:
        fl = Layout(dim, [([4, 8], [19, 23]), ([20, 0], [31, 15])], 3, P.nstate)
    # constant
    fbuf = _fine_field(fl, dim, lambda n, X: 1.0 + n)
    cbuf = np.full(cl.total, -7.0)
    oracle.level_average_down(P, fl, fbuf, cl, cbuf, ratio)
    for b in range(cl.nbox):
        v = cl.fab_view(cbuf, b)
        assert set(np.unique(v[0])) <= {-7.0, 1.0}
    # linear in the fine cell centres: coarse = its own centre value
    h = 1.0 / ratio
    fbuf = _fine_field(fl, dim, lambda n, X: 0.5 + sum((d + 1) * 0.3 * (X[d] + 0.5) * h for d in range(dim)))
    cbuf = np.full(cl.total, np.nan)
    oracle.level_average_down(P, fl, fbuf, cl, cbuf, ratio)
    covered = 0
    for b, (lo, hi) in enumerate(cl.boxes):
        v = cl.fab_view(cbuf, b)[0]
        idx = np.indices(v.shape)
        Xc = [idx[dim - 1 - d] + lo[d] - cl.ng for d in range(dim)]
        exact = 0.5 + sum((d + 1) * 0.3 * (Xc[d] + 0.5) for d in range(dim))
        m = ~np.isnan(v)
        covered += int(m.sum())
        if m.any():
            assert np.max(np.abs(v[m] - exact[m])) < 1e-13
    nfine = sum(int(np.prod([h - l + 1 for l, h in zip(lo, hi)])) for lo, hi in fl.boxes)
    assert covered == nfine // ratio ** dim
    # conservation under each fine box, random data
    rng = np.random.default_rng(3)
    fbuf = rng.uniform(0.5, 2.0, fl.total)
    cbuf = np.zeros(cl.total)
    oracle.level_average_down(P, fl, fbuf, cl, cbuf, ratio)
    fine_sum = sum(fl.fab_view(fbuf, b)[(slice(None),) + (slice(3, -3),) * dim].sum(axis=tuple(range(1, dim + 1)))
                   for b in range(fl.nbox))
    coarse_sum = sum(cl.fab_view(cbuf, b)[(slice(None),) + (slice(3, -3),) * dim].sum(axis=tuple(range(1, dim + 1)))
                     for b in range(cl.nbox))
    assert np.allclose(coarse_sum * ratio ** dim, fine_sum, rtol=1e-13)


def test_misaligned_fine_box_is_refused():
    P, cl, _ = _setup(2, 2)
    fl = Layout(2, [([3, 4], [10, 11])], 3, P.nstate)
    with pytest.raises(RuntimeError, match="failed: 7"):
        oracle.level_average_down(P, fl, np.zeros(fl.total), cl, np.zeros(cl.total), 2)


def _two_levels(ng=4):
    from compas_amd import make_geometry
    P = make_params(2, gamma=(1.4, 1.6))
    cg = make_geometry((16, 16), (0.0, 0.0), (1.0, 1.0), (0, 0), lo_bc=[2, 4], hi_bc=[2, 4])
    fg = make_geometry((32, 32), (0.0, 0.0), (1.0, 1.0), (0, 0), lo_bc=[2, 4], hi_bc=[2, 4])
    cl = Layout(2, split_domain((16, 16), 8), 3, P.nstate)
    fl = Layout(2, [([8, 8], [23, 23]), ([24, 8], [31, 23])], ng, P.nstate)
    return P, cg, fg, cl, fl


def _coarse_fill(cl, f):
    buf = np.zeros(cl.total)
    for b, (lo, hi) in enumerate(cl.boxes):
        v = cl.fab_view(buf, b)
        idx = np.indices(v.shape[1:])
        X = [idx[1 - d] + lo[d] - cl.ng for d in range(2)]
        for n in range(v.shape[0]):
            v[n] = f(n, X)
    return buf


def test_lincc_is_exact_for_linear_data_and_copies_fine_sources():
    P, cg, fg, cl, fl = _two_levels()
    lin = lambda n, X: 1.0 + n + (0.3 + 0.1 * n) * (X[0] + 0.5) + (0.2 - 0.05 * n) * (X[1] + 0.5)
    cbuf = _coarse_fill(cl, lin)
    fbuf = np.zeros(fl.total)
    for b in range(fl.nbox):          # fine valid cells: anything (only their ghosts are checked)
        fl.fab_view(fbuf, b)[...] = 100.0 + b
    oracle.level_fill_two_levels(P, fg, fl, fbuf, cg, cl, cbuf, 2)
    v = fl.fab_view(fbuf, 0)
    ng = fl.ng
    for n in (0, 4):
        for jj in range(v.shape[1]):
            for ii in range(v.shape[2]):
                i, j = 8 - ng + ii, 8 - ng + jj
                if 8 <= i <= 23 and 8 <= j <= 23:
                    continue
                if 24 <= i <= 31 and 8 <= j <= 23:
                    assert v[n, jj, ii] == 101.0          # a same-level source is copied
                    continue
                if not (0 <= j <= 31):
                    continue                             # slip walls mirror; checked below
                exact = 1.0 + n + (0.3 + 0.1 * n) * (i + 0.5) / 2 + (0.2 - 0.05 * n) * (j + 0.5) / 2
                assert abs(v[n, jj, ii] - exact) < 1e-13, (i, j, n)


def test_lincc_conserves_and_stays_bounded():
    P, cg, fg, cl, fl = _two_levels(ng=4)
    rng = np.random.default_rng(4)
    cbuf = _coarse_fill(cl, lambda n, X: rng.uniform(0.5, 2.0, X[0].shape))
    fbuf = np.zeros(fl.total)
    oracle.level_fill_two_levels(P, fg, fl, fbuf, cg, cl, cbuf, 2)
    v = fl.fab_view(fbuf, 0)                       # fine box (8..23)^2 with 4 ghosts: 2 coarse cells
    cv = {}
    for b, (lo, hi) in enumerate(cl.boxes):
        cf = cl.fab_view(cbuf, b)
        for j in range(lo[1], hi[1] + 1):
            for i in range(lo[0], hi[0] + 1):
                cv[(i, j)] = cf[:, j - lo[1] + 3, i - lo[0] + 3]
    checked = 0
    for jc in range(2, 14):
        for ic in range(2, 14):
            if 4 <= ic <= 11 and 4 <= jc <= 11:
                continue                             # under the fine box
            if ic >= 12 and 4 <= jc <= 11:
                continue                             # under the second fine box (copied)
            if not (2 <= ic <= 13 and 2 <= jc <= 13):
                continue
            ii, jj = 2 * ic - 8 + 4, 2 * jc - 8 + 4
            if not (0 <= ii and ii + 1 < v.shape[2] and 0 <= jj and jj + 1 < v.shape[1]):
                continue
            kids = v[:, jj:jj + 2, ii:ii + 2]
            assert np.allclose(kids.mean(axis=(1, 2)), cv[(ic, jc)], rtol=1e-14, atol=0)
            nb = np.array([cv[(ic + a, jc + b)] for a in (-1, 0, 1) for b in (-1, 0, 1)
                           if (ic + a, jc + b) in cv])
            assert np.all(kids.min(axis=(1, 2)) >= nb.min(axis=0) - 1e-14)
            assert np.all(kids.max(axis=(1, 2)) <= nb.max(axis=0) + 1e-14)
            checked += 1
    assert checked > 20


# ---- the stage-coupled step of a hierarchy ---------------------------------------------------------

def _hier(dim, fine_boxes, nc=16, seed=3):
    from compas_amd import cases
    h = cases.two_level_smooth(dim, nc, seed, fine_boxes)
    P = h["params"]
    lays = [Layout(dim, b, 3, P.nstate) for b in h["boxes"]]
    bufs = [lay.pack(U) for lay, U in zip(lays, h["fields"])]
    return h, P, lays, bufs


def test_amr_step_of_one_level_is_the_level_step():
    h, P, lays, bufs = _hier(2, [([8, 8], [15, 15])])
    ref = bufs[0].copy()
    oracle.level_rk3_step(P, h["geoms"][0], lays[0], ref, 1.0e-3)
    cm = oracle.amr_rk3_step(P, h["geoms"][:1], lays[:1], bufs[:1], 1.0e-3, 2)
    assert np.array_equal(lays[0].unpack(bufs[0], h["shapes"][0]), lays[0].unpack(ref, h["shapes"][0]))
    assert cm[0] > 0


def test_amr_step_with_a_covering_fine_level():
    """A fine level over the whole (periodic) domain takes no coarse data: it steps exactly as the
    fine level alone, and the coarse level ends as its average."""
    h, P, lays, bufs = _hier(2, [([0, 0], [15, 31]), ([16, 0], [31, 15]), ([16, 16], [31, 31])])
    alone = bufs[1].copy()
    dt = 2.0e-3
    oracle.amr_rk3_step(P, h["geoms"], lays, bufs, dt, 2)
    _single_step(P, h["geoms"][1], lays[1], alone, dt)
    fine = lays[1].unpack(bufs[1], h["shapes"][1])
    assert np.array_equal(fine, lays[1].unpack(alone, h["shapes"][1]))
    coarse = lays[0].unpack(bufs[0], h["shapes"][0])
    avg = fine.reshape(fine.shape[0], 16, 2, 16, 2).sum(axis=(2, 4)) * 0.25
    assert np.allclose(coarse, avg, rtol=1e-15, atol=0)


def _single_step(P, geom, lay, buf, dt):
    oracle.level_rk3_step(P, geom, lay, buf, dt)


@pytest.mark.parametrize("fine", [[([8, 8], [23, 23])], [([4, 8], [15, 19]), ([16, 4], [27, 23])],
                                  [([0, 8], [7, 23]), ([24, 8], [31, 15])]])
def test_amr_step_conserves_the_composite_totals(fine):
    """Face averaging makes the coarse-fine interface conservative: mass of each phase, momentum and
    energy summed over the coarse level (whose covered cells hold the fine means) stay put to
    round-off on a periodic domain. The last layout has fine boxes on the domain faces, so the
    interface crosses the periodic boundary."""
    h, P, lays, bufs = _hier(2, fine)
    c0 = lays[0].unpack(bufs[0], h["shapes"][0])
    oracle.level_average_down(P, lays[1], bufs[1], lays[0], bufs[0], 2)
    c0 = lays[0].unpack(bufs[0], h["shapes"][0])
    tot0 = c0.reshape(c0.shape[0], -1).sum(axis=1)
    for s in range(3):
        oracle.amr_rk3_step(P, h["geoms"], lays, bufs, 2.0e-3, 2)
    c1 = lays[0].unpack(bufs[0], h["shapes"][0])
    tot1 = c1.reshape(c1.shape[0], -1).sum(axis=1)
    scale = np.abs(c0).reshape(c0.shape[0], -1).sum(axis=1)
    for n in (0, 1, 2, 3, 4):
        assert abs(tot1[n] - tot0[n]) <= 1e-13 * scale[n], (n, tot1[n] - tot0[n], scale[n])
    assert np.abs(c1 - c0).max() > 1e-6


def test_amr_step_keeps_a_uniform_state():
    h, P, lays, bufs = _hier(2, [([8, 8], [23, 23])])
    for lay, buf in zip(lays, bufs):
        for b in range(lay.nbox):
            v = lay.fab_view(buf, b)
            for n in range(v.shape[0]):
                v[n] = h["fields"][0][n].mean()
    ref = [b.copy() for b in bufs]
    oracle.amr_rk3_step(P, h["geoms"], lays, bufs, 2.0e-3, 2)
    # (the stage combinations a U + (1 - a) U round: equal to a few ulps, not bit for bit)
    for lay, a, b in zip(lays, bufs, ref):
        for q in range(lay.nbox):
            sl = (slice(None),) + (slice(3, -3),) * 2
            x, y = lay.fab_view(a, q)[sl], lay.fab_view(b, q)[sl]
            assert np.all(np.abs(x - y) <= 4e-16 * np.abs(y)), np.abs(x - y).max()


def test_amr_compute_dt_is_the_smallest_level_estimate():
    from compas_amd import make_geometry
    g0 = make_geometry((16, 16), (0.0, 0.0), (1.0, 1.0), (1, 1))
    g1 = make_geometry((32, 32), (0.0, 0.0), (1.0, 1.0), (1, 1))
    cm = np.array([3.0, 2.0])
    est = [0.4 * (1 / 16) / 3.0, 0.4 * (1 / 32) / 2.0]
    assert oracle.amr_compute_dt([g0, g1], cm, 0.4, 1, 0.0) == min(est)
    assert oracle.amr_compute_dt([g0, g1], cm, 0.4, 1, 0.005) == min(min(est), 1.1 * 0.005)
    assert oracle.amr_compute_dt([g0, g1], cm, 0.4, 0, 0.005) == min(est)


# ---- a regridded level's data: old boxes where they hold the cell, else the coarse interpolation --

def test_fill_from_the_same_boxes_keeps_the_level_and_fills_its_ghosts():
    P, cg, fg, cl, fl = _two_levels()
    rng = np.random.default_rng(9)
    cbuf = rng.uniform(0.5, 2.0, cl.total)
    old = rng.uniform(0.5, 2.0, fl.total)
    ref = old.copy()
    oracle.level_fill_two_levels(P, fg, fl, ref, cg, cl, cbuf, 2)
    new = np.zeros(fl.total)
    oracle.level_fill_from(P, fg, fl, new, fl, old, cg, cl, cbuf, 2)
    assert np.array_equal(new, ref)


def test_fill_from_the_coarse_level_alone_is_conservative_and_linear_exact():
    P, cg, fg, cl, fl = _two_levels()
    lin = lambda n, X: 1.0 + n + (0.3 + 0.1 * n) * (X[0] + 0.5) + (0.2 - 0.05 * n) * (X[1] + 0.5)
    cbuf = _coarse_fill(cl, lin)
    new = np.zeros(fl.total)
    oracle.level_fill_from(P, fg, fl, new, None, None, cg, cl, cbuf, 2)
    v = fl.fab_view(new, 0)
    g = fl.ng
    for jj in range(v.shape[1]):
        for ii in range(v.shape[2]):
            i, j = 8 + ii - g, 8 + jj - g
            if not (0 <= j <= 31):
                continue                             # slip walls mirror there
            exact = 1.0 + (0.3) * (i + 0.5) / 2 + 0.2 * (j + 0.5) / 2
            assert abs(v[0, jj, ii] - exact) < 1e-13
    rng = np.random.default_rng(2)
    cbuf = _coarse_fill(cl, lambda n, X: rng.uniform(0.5, 2.0, X[0].shape))
    oracle.level_fill_from(P, fg, fl, new, None, None, cg, cl, cbuf, 2)
    back = cbuf.copy()
    oracle.level_average_down(P, fl, new, cl, back, 2)
    assert np.allclose(back, cbuf, rtol=1e-14, atol=0)


def test_fill_from_partly_overlapping_old_boxes_copies_where_they_hold_the_cell():
    P, cg, fg, cl, fl = _two_levels()
    old_l = Layout(2, [([12, 12], [19, 27])], 3, P.nstate)
    rng = np.random.default_rng(5)
    cbuf = rng.uniform(0.5, 2.0, cl.total)
    old = rng.uniform(0.5, 2.0, old_l.total)
    new = np.zeros(fl.total)
    oracle.level_fill_from(P, fg, fl, new, old_l, old, cg, cl, cbuf, 2)
    alone = np.zeros(fl.total)
    oracle.level_fill_from(P, fg, fl, alone, None, None, cg, cl, cbuf, 2)
    v, a, o = fl.fab_view(new, 0), fl.fab_view(alone, 0), old_l.fab_view(old, 0)
    g = fl.ng
    for jj in range(v.shape[1]):
        for ii in range(v.shape[2]):
            i, j = 8 - g + ii, 8 - g + jj
            if 12 <= i <= 19 and 12 <= j <= 27:
                assert np.array_equal(v[:, jj, ii], o[:, j - 12 + 3, i - 12 + 3])
            elif 0 <= j <= 31:
                assert np.array_equal(v[:, jj, ii], a[:, jj, ii])
