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
