"""Pins of the oracle's refinement tagging (ErrorEst with include/Tagging.H's criteria; CPU only):

- state_grad on a linear field tags every cell exactly when |slope| dx >= the threshold (the
  one-sided differences equal |slope| dx; equality tags), in each direction, 3D included;
- value_less / value_greater / value_range are strict and match numpy on the same values;
- the derived variables: rho is the sum of the partial densities, u = rho u / rho, P is the EOS
  pressure (prim_to_cons of a known pressure field gives it back);
- the criteria combine by OR; the tags follow the box split (box by box, x fastest).
"""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, cases, make_geometry, make_params, split_domain


def _level(dim, n, W, P, mg=None, ng=3):
    U = cases.prim_to_cons(P, W) if not P.nphase else cases.np_prim_to_cons(P, W)
    shape = (n,) * dim
    G = make_geometry(shape, (0.0,) * dim, (1.0,) * dim, (0,) * dim, lo_bc=[2] * dim, hi_bc=[2] * dim)
    boxes = split_domain(shape, mg or n)
    lay = Layout(dim, boxes, ng, P.nstate)
    buf = lay.pack(U)
    oracle.level_fill_ghosts(P, G, lay, buf)
    return lay, buf, boxes, U


def _global(lay, boxes, tags, shape):
    out = np.zeros(shape, dtype=np.uint8)
    q = 0
    for lo, hi in boxes:
        ext = [h - l + 1 for l, h in zip(lo, hi)]
        n = int(np.prod(ext))
        sl = tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(len(lo))))
        out[sl] = tags[q:q + n].reshape(tuple(reversed(ext)))
        q += n
    return out


@pytest.mark.parametrize("dim,axis", [(2, 0), (2, 1), (3, 2)])
def test_grad_on_a_linear_field(dim, axis):
    n = 16
    P = make_params(dim, gamma=(1.4, 1.6))
    W = np.empty((P.nstate,) + (n,) * dim)
    X = np.indices((n,) * dim)[dim - 1 - axis].astype(float)   # index along `axis`
    a1 = 0.2 + 0.03 * X                                          # step 0.03 per cell
    W[0], W[1] = a1 * 1.0, (1 - a1) * 2.0
    W[2:2 + dim] = 0.0
    W[2 + dim] = 1.0
    W[3 + dim] = a1
    lay, buf, boxes, _ = _level(dim, n, W, P, mg=8)
    comp = 3 + dim
    # outflow ghosts repeat the edge value: the edge cells still see the 0.03 step on one side
    hit = oracle.level_tag(P, lay, buf, [(comp, "grad", 0.03 - 1e-12)])
    assert hit.all()
    miss = oracle.level_tag(P, lay, buf, [(comp, "grad", 0.03 + 1e-12)])
    assert not miss.any()


def test_value_criteria_are_strict_and_or_combined():
    dim, n = 2, 12
    P = make_params(dim, gamma=(1.4, 1.6))
    rng = np.random.default_rng(1)
    a1 = rng.choice([0.1, 0.25, 0.5, 0.75, 0.9], size=(n, n))
    W = np.empty((P.nstate, n, n))
    W[0], W[1] = a1, (1 - a1) * 3.0
    W[2:4] = 0.0
    W[4] = 1.0
    W[5] = a1
    lay, buf, boxes, U = _level(dim, n, W, P, mg=5)
    g = lambda crit: _global(lay, boxes, oracle.level_tag(P, lay, buf, crit), (n, n))
    assert np.array_equal(g([(5, "value_less", 0.25)]), (a1 < 0.25).astype(np.uint8))
    assert np.array_equal(g([(5, "value_greater", 0.75)]), (a1 > 0.75).astype(np.uint8))
    assert np.array_equal(g([(5, "value_range", 0.25, 0.75)]), ((a1 > 0.25) & (a1 < 0.75)).astype(np.uint8))
    both = g([(5, "value_less", 0.25), (5, "value_greater", 0.75)])
    assert np.array_equal(both, ((a1 < 0.25) | (a1 > 0.75)).astype(np.uint8))


@pytest.mark.parametrize("nphase", [0, 3])
def test_derived_variables(nphase):
    dim, n = 2, 10
    rng = np.random.default_rng(2)
    if nphase:
        P = make_params(dim, gamma=(1.4, 1.67, 1.09), nphase=3)
        a = rng.dirichlet([2, 2, 2], size=(n, n)).transpose(2, 0, 1)
        W = np.empty((P.nstate, n, n))
        W[0:3] = a * np.array([1.0, 0.14, 5.0])[:, None, None]
        u, v = rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, n))
        W[3], W[4] = u, v
        p = rng.uniform(0.5, 2.0, (n, n))
        W[5] = p
        W[6:9] = a
        rho = W[0] + W[1] + W[2]
    else:
        P = make_params(dim, gamma=(1.4, 1.6))
        a1 = rng.uniform(0.05, 0.95, (n, n))
        W = np.empty((P.nstate, n, n))
        W[0], W[1] = a1 * 1.0, (1 - a1) * 2.0
        u, v = rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, n))
        W[2], W[3] = u, v
        p = rng.uniform(0.5, 2.0, (n, n))
        W[4] = p
        W[5] = a1
        rho = W[0] + W[1]
    lay, buf, boxes, U = _level(dim, n, W, P, mg=5)
    g = lambda crit: _global(lay, boxes, oracle.level_tag(P, lay, buf, crit), (n, n))
    for name, field in (("rho", rho), ("P", p), ("u", u), ("v", v)):
        med = np.median(field)
        # the threshold sits between distinct values, so round-off cannot move a cell across it
        vals = np.sort(field.ravel())
        k = np.searchsorted(vals, med)
        thr = 0.5 * (vals[k - 1] + vals[k])
        assert np.array_equal(g([(name, "value_greater", thr)]), (field > thr).astype(np.uint8)), name
