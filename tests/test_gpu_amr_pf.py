"""GPU parity of the Phase-Field step on a hierarchy (cmp_amr_pf_*, through the C ABI) against the
oracle on the same hierarchies: two and three properly nested levels, the interfaces crossing
coarse-fine boundaries, two- and three-phase (three gases with separate interfaces), ACDI with the
explicit and implicit RC and CAC-Adv with the composite implicit solve, 2D and 3D.

Bars: Grad_Q of every level <= 1e-12 of its largest value with the explicit RC and <= 1e-9 with the
implicit RC (both sides solve the same composite system to the rounding floor with sums in other
orders; an unconverged or wrong solve is off by O(1)), equal mechanism and Phase-Field sub-step
counts; the whole step <= 1e-12 (explicit) / 1e-10 (implicit) per variable max relative difference;
flow + Phase-Field steps <= 1e-9 in L1 with the composite totals conserved to 1e-13.
"""
import numpy as np
import pytest

import oracle
from compas_amd import Amr, Layout, Level, cases, make_pf_params
from tests._util import l1rel_per_var, maxrel_per_var

pytestmark = pytest.mark.gpu

NG = 3
CF_CROSSING = [([4, 4], [15, 27]), ([16, 4], [19, 11])]
THREE = [[([4, 4], [27, 27])], [([14, 14], [35, 43]), ([36, 14], [49, 27])]]


def _case(name):
    if name == "2d":
        return cases.interface_hierarchy(2, 16, [CF_CROSSING], width=2.0)
    if name == "2d-neumann":
        return cases.interface_hierarchy(2, 16, [CF_CROSSING], width=2.0, periodic=False)
    if name == "2d-3phase":
        return cases.interface_hierarchy(2, 16, [CF_CROSSING], width=2.0, nphase=3, rings=True)
    if name == "2d-3phase-3lev":
        return cases.interface_hierarchy(2, 16, THREE, width=2.0, nphase=3, rings=True)
    if name == "2d-3lev":
        return cases.interface_hierarchy(2, 16, THREE, width=2.0)
    if name == "3d":
        return cases.interface_hierarchy(3, 8, [[([2, 2, 2], [7, 13, 13])]], width=2.0)
    raise ValueError(name)


def _gpu_hier(gpu, c):
    import torch
    P = c["params"]
    lays = [Layout(P.dim, b, NG, P.nstate) for b in c["boxes"]]
    bufs = [lay.pack(f) for lay, f in zip(lays, c["fields"])]
    levels = [Level(P, g, lay) for g, lay in zip(c["geoms"], lays)]
    amr = Amr(levels, 2)
    U = [torch.from_numpy(b.copy()).to(gpu) for b in bufs]
    return P, lays, bufs, levels, amr, U


def _close(levels, amr):
    amr.close()
    for lv in levels:
        lv.close()


CASES = [("2d", "ACDI", 1), ("2d", "ACDI", 0), ("2d", "CAC-Adv", 0), ("2d-neumann", "CAC-Adv", 0),
         ("2d-3lev", "CAC-Adv", 0), ("2d-3lev", "ACDI", 1), ("2d-3phase", "ACDI", 1), ("2d-3phase", "ACDI", 0),
         ("2d-3phase-3lev", "CAC-Adv", 0), ("2d-3phase-3lev", "ACDI", 1), ("3d", "ACDI", 1), ("3d", "CAC-Adv", 0)]


@pytest.mark.parametrize("name,mech,rc", CASES)
def test_hierarchy_gradq_matches_oracle(gpu, name, mech, rc):
    import torch
    from compas_amd import FaceLayout
    c = _case(name)
    P, lays, bufs, levels, amr, U = _gpu_hier(gpu, c)
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, eta_multiplier=0.5, cfl=0.3)
    dt = 0.5 / c["shapes"][0][0]
    ref, ref_info = oracle.amr_pf_gradq_step(P, pf, c["geoms"], lays, [b.copy() for b in bufs], dt, 2)
    npf = P.nphase if P.nphase else 1
    gq = [torch.full((FaceLayout(P.dim, lay.boxes, npf).total,), float("nan"), dtype=torch.float64, device=gpu)
          for lay in lays]
    print("hierarchy Grad_Q %s %s rc=%d: %d levels" % (name, mech, rc, len(lays)), flush=True)
    amr.pf_setup(pf)
    gm, info = amr.PhaseField_GradQPF(U, dt, gq)
    assert amr.check_errors() == 0
    got = [g.cpu().numpy() for g in gq]
    its = amr.solver_iterations()
    _close(levels, amr)
    assert info[3] == ref_info[3] and info[3] >= 1
    for q in range(3):
        assert abs(info[q] - ref_info[q]) <= 1e-15 * abs(ref_info[q])
    if rc == 0:
        print("composite CG iterations: GPU %d (oracle Jacobi-PCG %d)" % (its, ref_info[4]), flush=True)
    scale = max(np.max(np.abs(r)) for r in ref)
    assert scale > 0.0
    tol = 1e-12 if rc == 1 else 1e-9
    for l in range(len(lays)):
        assert np.all(np.isfinite(got[l]))
        err = np.max(np.abs(got[l] - ref[l]))
        assert err <= tol * scale, (l, err / scale)
        assert abs(gm[l] - np.max(np.abs(ref[l]))) <= tol * scale


@pytest.mark.parametrize("name,mech,rc", [("2d", "ACDI", 1), ("2d", "CAC-Adv", 0), ("2d-3lev", "ACDI", 1),
                                          ("2d-3phase", "ACDI", 1), ("2d-3phase-3lev", "CAC-Adv", 0),
                                          ("3d", "ACDI", 1)])
def test_hierarchy_pf_step_matches_oracle(gpu, name, mech, rc):
    import torch
    c = _case(name)
    P, lays, bufs, levels, amr, U = _gpu_hier(gpu, c)
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, eta_multiplier=0.5, cfl=0.3)
    dt = 0.5 / c["shapes"][0][0]
    ref_bufs = [b.copy() for b in bufs]
    ref_info = oracle.amr_pf_evolve(P, pf, c["geoms"], lays, ref_bufs, dt, 2)
    amr.pf_setup(pf)
    U1 = [torch.zeros_like(u) for u in U]
    U2 = [torch.zeros_like(u) for u in U]
    dtd = torch.full((1,), dt, dtype=torch.float64, device=gpu)
    print("hierarchy Phase-Field step %s %s rc=%d" % (name, mech, rc), flush=True)
    info = amr.PhaseField_Evolve(U, U1, U2, dtd)
    assert amr.check_errors() == 0
    _close(levels, amr)
    assert info[3] == ref_info[3] and info[5] == ref_info[5] and info[5] >= 1
    tol = 1e-12 if rc == 1 else 1e-10
    for l, (lay, shp) in enumerate(zip(lays, c["shapes"])):
        got = lay.unpack(U[l].cpu().numpy(), shp)
        ref = lay.unpack(ref_bufs[l], shp)
        m = np.zeros(shp, dtype=bool)
        for lo, hi in lay.boxes:
            m[tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(P.dim)))] = True
        for q in range(P.nstate):
            scale = np.max(np.abs(ref[q][m]))
            assert np.max(np.abs(got[q][m] - ref[q][m])) <= tol * scale, (l, q)


@pytest.mark.parametrize("mech,rc", [("CAC-Adv", 0), ("ACDI", 1)])
def test_hierarchy_advance_with_pf_matches_oracle(gpu, mech, rc):
    """Flow + Phase-Field steps of a two-level hierarchy (graph-replayed flow part): dt history to
    1e-12, L1 <= 1e-9 on every level, composite totals conserved."""
    import torch
    c = _case("2d")
    P, lays, bufs, levels, amr, U = _gpu_hier(gpu, c)
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, eta_multiplier=0.5, cfl=0.3)
    ref_bufs = [b.copy() for b in bufs]
    nsteps = 5
    hist_ref = oracle.amr_advance_pf(P, pf, c["geoms"], lays, ref_bufs, nsteps, 0.4, 2)
    amr.pf_setup(pf)
    U1 = [torch.zeros_like(u) for u in U]
    U2 = [torch.zeros_like(u) for u in U]
    dt = torch.zeros(1, dtype=torch.float64, device=gpu)
    cm = torch.zeros(len(U), dtype=torch.float64, device=gpu)
    hist = torch.zeros(nsteps, dtype=torch.float64, device=gpu)
    amr.advance(U, U1, U2, nsteps, 0.4, dt, cm, dt_hist=hist, use_graph=True)
    assert amr.check_errors() == 0
    _close(levels, amr)
    assert np.max(np.abs(hist.cpu().numpy() - hist_ref) / hist_ref) <= 1e-12
    tot0 = c["fields"][0].reshape(P.nstate, -1).sum(axis=1)
    for l, (lay, shp) in enumerate(zip(lays, c["shapes"])):
        got = lay.unpack(U[l].cpu().numpy(), shp)
        ref = lay.unpack(ref_bufs[l], shp)
        m = np.zeros(shp, dtype=bool)
        for lo, hi in lay.boxes:
            m[tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(P.dim)))] = True
        l1 = l1rel_per_var(got[:, m], ref[:, m])
        assert l1.max() <= 1e-9, (l, l1)
        if l == 0:
            tot = got.reshape(P.nstate, -1).sum(axis=1)
            assert np.all(np.abs(tot - tot0) <= 1e-13 * np.abs(got).reshape(P.nstate, -1).sum(axis=1))


def test_step_regrid_step_with_phase_field(gpu):
    """Flow + CAC Phase-Field steps of a two-level hierarchy with a regrid every 2 steps inside the
    run (RegridHierarchy: GPU tags of the volume fraction's gradient, host clustering, the fine level
    made again from its old data and the coarse level, average-down, the hierarchy and its
    Phase-Field buffers rebuilt): no device error, the fine boxes move to the tagged interface,
    composite totals conserved to 1e-12 across the regrids, u and p kept uniform."""
    import torch
    from compas_amd.amr_run import RegridHierarchy
    c = _case("2d")
    P, lays, bufs, levels, amr, U = _gpu_hier(gpu, c)
    amr.close()
    pf = make_pf_params(mechanism="CAC-Adv", explicit_rc=0, eta_multiplier=0.5, cfl=0.3)
    h = RegridHierarchy(P, c["geoms"], levels, U, 2, criteria=[(5, "grad", 0.05)], blocking_factor=8,
                        max_grid_size=16, pf=pf)
    tot0 = c["fields"][0].reshape(P.nstate, -1).sum(axis=1)
    h.advance(6, 0.4, regrid_int=2)
    assert h.check_errors() == 0
    assert h.stats["regrids"] == 2
    boxes = h.levels[1].layout.boxes
    assert boxes != [(list(lo), list(hi)) for lo, hi in c["boxes"][1]]
    U0 = lays[0].unpack(h.Us[0].cpu().numpy(), c["shapes"][0])
    tot = U0.reshape(P.nstate, -1).sum(axis=1)
    assert np.all(np.abs(tot - tot0) <= 1e-12 * np.abs(U0).reshape(P.nstate, -1).sum(axis=1)), (tot - tot0)
    lay1 = h.levels[1].layout
    for lay, Ul, shp in ((lays[0], h.Us[0], c["shapes"][0]), (lay1, h.Us[1], c["shapes"][1])):
        W = lay.unpack(Ul.cpu().numpy(), shp)
        m = np.zeros(shp, dtype=bool)
        for lo, hi in lay.boxes:
            m[tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(P.dim)))] = True
        rho = W[0][m] + W[1][m]
        assert np.max(np.abs(W[2][m] / rho - 1.0)) < 1e-10 and np.max(np.abs(W[3][m] / rho - 0.5)) < 1e-10
    h.close()
