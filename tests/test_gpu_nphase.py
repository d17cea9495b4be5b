"""GPU parity of the N-phase model (nphase = 2 or 3) through the C ABI against the oracle on the
same seeded inputs: the per-box RHS of every built scheme x Riemann solver x K mode, the ghost fill
with reflecting walls, one RK3 step and 50 steps of a level. Bars as for the two-phase model:
per-variable max relative difference <= 1e-12 for one RHS or one step, <= 1e-9 in L1 after the
steps; the ghost fill bit-exact; the volume fractions keep summing to 1 (the consistent limiter);
face arrays of the N-phase model are refused.
"""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, Level, cases, make_geometry, rhs_box
from compas_amd.capi import CmpError
from tests._util import l1rel_per_var, maxrel_per_var

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12


def _dev(torch, a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


CASES = [(2, 3, "WENO5", "HLLC", 0), (2, 3, "WENO5", "HLL", 0), (2, 3, "MUSCL", "HLLC", 0),
         (2, 3, "GODUNOV", "HLL", 0), (2, 3, "WENO5", "HLLC", 1), (2, 2, "WENO5", "HLLC", 0),
         (2, 2, "MUSCL", "HLL", 1), (3, 3, "WENO5", "HLLC", 0), (3, 3, "MUSCL", "HLL", 1),
         (3, 2, "WENO5", "HLLC", 0)]


@pytest.mark.parametrize("dim,nphase,scheme,riemann,source_term", CASES)
def test_nphase_rhs_matches_oracle(gpu, dim, nphase, scheme, riemann, source_term):
    import torch
    n = 24 if dim == 2 else 12
    c = cases.nphase_smooth(dim, n, seed=5 + dim, nphase=nphase, scheme=scheme, riemann=riemann,
                            pinf=(0.0, 2.0, 0.0), source_term=source_term)
    P, G = c["params"], c["geom"]
    ng = 3
    lay = Layout(dim, c["boxes"], ng, P.nstate)
    buf = lay.pack(c["U"])
    oracle.level_fill_ghosts(P, G, lay, buf)
    fab = lay.fab_view(buf, 0).copy()
    ref, ref_cmax = oracle.rhs_box(P, G, c["boxes"][0], ng, fab)
    U = _dev(torch, fab.ravel(), gpu)
    dUdt = torch.zeros(ref.size, dtype=torch.float64, device=gpu)
    cmax = torch.zeros(1, dtype=torch.float64, device=gpu)
    rhs_box(P, G, c["boxes"][0], ng, U, dUdt, cmax)
    torch.cuda.synchronize()
    got = dUdt.cpu().numpy().reshape(ref.shape)
    rel = maxrel_per_var(got, ref)
    assert rel.max() <= TOL_STEP, rel
    assert abs(cmax.item() - ref_cmax) <= 1e-14 * ref_cmax
    # face arrays are not built for this model
    fl = [torch.zeros(8, dtype=torch.float64, device=gpu)] + [None] * 2
    with pytest.raises(CmpError):
        rhs_box(P, G, c["boxes"][0], ng, U, dUdt, cmax, fl, [None] * 3)


@pytest.mark.parametrize("dim,bc", [(2, (3, 5)), (2, (4, 2)), (3, (5, 3))])
def test_nphase_fill_bit_exact(gpu, dim, bc):
    import torch
    n = 24 if dim == 2 else 12
    c = cases.nphase_smooth(dim, n, seed=2, periodic=False, max_grid_size=n // 3)
    G = make_geometry((n,) * dim, (0.0,) * dim, (1.0,) * dim, (0,) * dim, lo_bc=[bc[0]] * dim,
                      hi_bc=[bc[1]] * dim)
    lay = Layout(dim, c["boxes"], 3, c["params"].nstate)
    buf = lay.pack(c["U"])
    ref = buf.copy()
    oracle.level_fill_ghosts(c["params"], G, lay, ref)
    U = _dev(torch, buf, gpu)
    lev = Level(c["params"], G, lay)
    lev.FillPatch(U)
    assert lev.check_errors() == 0
    assert np.array_equal(U.cpu().numpy(), ref)
    lev.close()


def _advance(c, lay, nsteps, gpu):
    import torch
    Un = _dev(torch, lay.pack(c["U"]), gpu)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dt = torch.zeros(1, dtype=torch.float64, device=gpu)
    cm = torch.zeros(1, dtype=torch.float64, device=gpu)
    hist = torch.zeros(max(1, nsteps), dtype=torch.float64, device=gpu)
    lev = Level(c["params"], c["geom"], lay)
    lev.advance(Un, U1, U2, nsteps, c["cfl"], dt, cm, hist, use_graph=True)
    assert lev.check_errors() == 0
    lev.close()
    return lay.unpack(Un.cpu().numpy(), c["shape"]), hist.cpu().numpy()[:nsteps]


@pytest.mark.parametrize("dim,scheme,riemann", [(2, "WENO5", "HLLC"), (2, "MUSCL", "HLL"), (3, "WENO5", "HLLC")])
def test_nphase_rk3_step_matches_oracle(gpu, dim, scheme, riemann):
    n = 32 if dim == 2 else 16
    c = cases.nphase_smooth(dim, n, seed=8, scheme=scheme, riemann=riemann, max_grid_size=n // 2)
    P = c["params"]
    lay = Layout(dim, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    hist_ref = oracle.level_advance(P, c["geom"], lay, buf, 1, c["cfl"])
    ref = lay.unpack(buf, c["shape"])
    got, hist = _advance(c, lay, 1, gpu)
    assert abs(hist[0] - hist_ref[0]) <= 1e-14 * hist_ref[0]
    rel = maxrel_per_var(got, ref)
    assert rel.max() <= TOL_STEP, rel


def test_nphase_50_steps_l1(gpu):
    """50 graph-replayed steps of a 2D three-phase state (64^2 in four boxes): L1 <= 1e-9 per
    variable, equal dt history, totals conserved, volume fractions summing to 1."""
    c = cases.nphase_smooth(2, 64, seed=12, max_grid_size=32)
    P = c["params"]
    lay = Layout(2, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    hist_ref = oracle.level_advance(P, c["geom"], lay, buf, 50, c["cfl"])
    ref = lay.unpack(buf, c["shape"])
    got, hist = _advance(c, lay, 50, gpu)
    assert np.max(np.abs(hist - hist_ref) / hist_ref) <= 1e-12
    l1 = l1rel_per_var(got, ref)
    assert l1.max() <= 1e-9, l1
    for q in range(P.nphase + 3):
        assert abs(got[q].sum() - c["U"][q].sum()) <= 1e-12 * np.abs(c["U"][q]).sum()
    assert np.max(np.abs(got[6:9].sum(axis=0) - 1.0)) <= 1e-12
