"""Pins of the oracle's Phase-Field step on a hierarchy (PhaseField_Evolve over finest_level + 1
levels without subcycling; CPU only):

- one level is the single-level Phase-Field step (mechanism, Grad_Q and sub-steps);
- a finer level over the whole domain gives, on that level, the step of the level alone;
- a whole step conserves the composite totals (level 0 after AverageDown) to round-off, keeps a
  uniform u and p uniform on every level while the interface moves, and keeps sum_k alpha_k = 1 for
  three phases, with the interface crossing a coarse-fine boundary;
- the composite implicit solve satisfies its own equation on every composite cell, with the face
  coefficients evaluated independently, and flux continuity across the coarse-fine faces.
"""
import math

import numpy as np
import pytest

import oracle
from compas_amd import FaceLayout, Layout, cases, make_pf_params

NG = 3


def _hier(c):
    P = c["params"]
    lays = [Layout(P.dim, b, NG, P.nstate) for b in c["boxes"]]
    bufs = [lay.pack(f) for lay, f in zip(lays, c["fields"])]
    return P, c["geoms"], lays, bufs


def _uv_p(P, U):
    """u (all components) and p of a field [n, ...] (ideal gases, the two- or N-phase state)."""
    dim = P.dim
    npf = P.nphase if P.nphase else 2
    rho = sum(U[k] for k in range(npf))
    mom0 = npf
    u = [U[mom0 + m] / rho for m in range(dim)]
    rhoE = U[mom0 + dim]
    rhoe = rhoE - 0.5 * rho * sum(x * x for x in u)
    if P.nphase:
        alpha = [U[P.nphase + dim + 1 + k] for k in range(P.nphase)]
    else:
        alpha = [U[3 + dim], 1.0 - U[3 + dim]]
    A = sum(alpha[k] * P.eos_A[k] for k in range(len(alpha)))
    return u, rhoe / A


CF_CROSSING = [([4, 4], [15, 27]), ([16, 4], [19, 11])]      # fine boxes: the interface leaves them


@pytest.mark.parametrize("mech,rc", [("ACDI", 1), ("ACDI", 0), ("CAC-Adv", 0)])
def test_one_level_is_the_level_step(mech, rc):
    c = cases.interface_hierarchy(2, 24, [], width=2.5, max_grid_size=12)
    P, G, lays, bufs = _hier(c)
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, eta_multiplier=0.5, cfl=0.3)
    dt = 0.3 / 24
    flay = FaceLayout(2, lays[0].boxes)
    ref, _, ref_info = oracle.level_pf_gradq_step(P, pf, G[0], lays[0], bufs[0].copy(), dt, flay.total)
    gqs, info = oracle.amr_pf_gradq_step(P, pf, G, lays, [bufs[0].copy()], dt, 2)
    assert info[3] == ref_info[3] and info[3] >= 1
    assert info[:3] == ref_info[:3]
    tol = 0.0 if rc == 1 else 1e-9
    assert np.max(np.abs(gqs[0] - ref)) <= tol * np.max(np.abs(ref)), np.max(np.abs(gqs[0] - ref))
    a, b = bufs[0].copy(), bufs[0].copy()
    ref_info = oracle.level_pf_evolve(P, pf, G[0], lays[0], a, dt)
    info = oracle.amr_pf_evolve(P, pf, G, lays, [b], dt, 2)
    assert info[5] == ref_info[5] and info[3] == ref_info[3]
    Ua, Ub = lays[0].unpack(a, c["shapes"][0]), lays[0].unpack(b, c["shapes"][0])
    scale = np.abs(Ua).reshape(P.nstate, -1).max(axis=1)
    err = np.abs(Ua - Ub).reshape(P.nstate, -1).max(axis=1) / scale
    assert err.max() <= (1e-14 if rc == 1 else 1e-10), err


@pytest.mark.parametrize("mech,rc", [("ACDI", 1), ("CAC-Adv", 0)])
def test_fine_level_over_the_whole_domain_is_that_level_alone(mech, rc):
    c = cases.interface_hierarchy(2, 12, [[([0, 0], [11, 23]), ([12, 0], [23, 23])]], width=2.5)
    P, G, lays, bufs = _hier(c)
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, eta_multiplier=0.5, cfl=0.3)
    dt = 0.25 / 24
    alone = bufs[1].copy()
    ref_info = oracle.level_pf_evolve(P, pf, G[1], lays[1], alone, dt)
    info = oracle.amr_pf_evolve(P, pf, G, lays, bufs, dt, 2)
    assert info[3] == ref_info[3] and info[5] == ref_info[5]
    Ua, Ub = lays[1].unpack(alone, c["shapes"][1]), lays[1].unpack(bufs[1], c["shapes"][1])
    scale = np.abs(Ua).reshape(P.nstate, -1).max(axis=1)
    err = np.abs(Ua - Ub).reshape(P.nstate, -1).max(axis=1) / scale
    # CAC's level sums are taken over the averaged-down coarse level (another order of additions)
    assert err.max() <= (1e-14 if rc == 1 else 1e-10), err


@pytest.mark.parametrize("dim,mech,rc,nphase", [(2, "ACDI", 1, 0), (2, "ACDI", 0, 0), (2, "CAC-Adv", 0, 0),
                                                (2, "CAC-Adv", 0, 3), (3, "ACDI", 1, 0)])
def test_hierarchy_step_conserves_and_keeps_u_p_uniform(dim, mech, rc, nphase):
    if dim == 2:
        c = cases.interface_hierarchy(2, 16, [CF_CROSSING], width=2.0, nphase=nphase)
    else:
        c = cases.interface_hierarchy(3, 8, [[([2, 2, 2], [7, 13, 13])]], width=2.0)
    P, G, lays, bufs = _hier(c)
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, eta_multiplier=0.5, cfl=0.3)
    n0 = c["shapes"][0][0]
    dt = 0.5 / n0
    tot0 = lays[0].unpack(bufs[0], c["shapes"][0]).reshape(P.nstate, -1).sum(axis=1)
    info = oracle.amr_pf_evolve(P, pf, G, lays, bufs, dt, 2)
    assert info[3] >= 1 and info[5] >= 1 and info[4] > 0.0
    U0 = lays[0].unpack(bufs[0], c["shapes"][0])
    tot = U0.reshape(P.nstate, -1).sum(axis=1)
    mag = np.abs(U0).reshape(P.nstate, -1).sum(axis=1)
    assert np.all(np.abs(tot - tot0) <= 1e-13 * mag), np.abs(tot - tot0) / mag
    moved = 0.0
    for lay, buf, shp, f0 in zip(lays, bufs, c["shapes"], c["fields"]):
        U = lay.unpack(buf, shp)
        # the valid cells of the level's boxes
        m = np.zeros(shp, dtype=bool)
        for lo, hi in lay.boxes:
            m[tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(dim)))] = True
        u, p = _uv_p(P, U)
        for k in range(dim):
            assert np.max(np.abs(u[k][m] - (1.0, 0.5, -0.25)[k])) <= 1e-12
        assert np.max(np.abs(p[m] - 1.0)) <= 1e-12
        ia = (P.nphase + dim + 1) if P.nphase else 3 + dim
        moved = max(moved, float(np.max(np.abs(U[ia][m] - f0[ia][m]))))
        if P.nphase:
            s = sum(U[P.nphase + dim + 1 + k] for k in range(P.nphase))
            assert np.max(np.abs(s[m] - 1.0)) <= 1e-12
    assert moved > 1e-6


def _global_faces(flay, buf, shape, d, lvl_boxes):
    """Face values of direction d of a level's face buffer into a global nodal array (NaN where the
    level has no face)."""
    ny, nx = shape
    out = np.full((ny + (d == 1), nx + (d == 0)), np.nan)
    for b, (lo, hi) in enumerate(lvl_boxes):
        v = flay.face_view(buf, b, d)[0]
        out[lo[1]:lo[1] + v.shape[0], lo[0]:lo[0] + v.shape[1]] = v
    return out


def test_composite_solve_satisfies_its_equation():
    """The composite implicit solve through the CAC step's Grad_Q: with the mechanism switched off
    (a zero-mobility interface is not possible, so the equation is checked on the solve itself):
    on every composite cell the fluxes b Grad_Q through its faces (a coarse cell next to the finer
    level through the fine faces on its face) sum to V (rhs - mean); the fine faces of a coarse-fine
    face carry the two-point difference across the interface; Neumann faces carry zero."""
    nc = 16
    c = cases.interface_hierarchy(2, nc, [CF_CROSSING], width=2.0, periodic=False)
    P, G, lays, bufs = _hier(c)
    pf = make_pf_params(mechanism="CAC-Adv", explicit_rc=0, eta_multiplier=0.5, eta_factor=2.0)
    alays = [Layout(2, lay.boxes, 1, 1) for lay in lays]
    # alpha of every level; right-hand side random on every level (covered cells too)
    rng = np.random.default_rng(5)
    alpha = [al.pack(f[5:6]) for al, f in zip(alays, c["fields"])]
    rhs = [al.pack(rng.uniform(-1, 1, (1,) + shp)) for al, shp in zip(alays, c["shapes"])]
    eta = 0.5 / (2 * nc)
    gq, bf, it = oracle.amr_pf_gradq_implicit(P, pf, G, alays, alpha, rhs, eta, 2)
    assert it >= 1
    flays = [FaceLayout(2, al.boxes) for al in alays]
    shp0, shp1 = c["shapes"]
    h0, h1 = 1.0 / nc, 1.0 / (2 * nc)
    gx = [_global_faces(fl, g, s, 0, al.boxes) for fl, g, s, al in zip(flays, gq, c["shapes"], alays)]
    gy = [_global_faces(fl, g, s, 1, al.boxes) for fl, g, s, al in zip(flays, gq, c["shapes"], alays)]
    bx = [_global_faces(fl, g, s, 0, al.boxes) for fl, g, s, al in zip(flays, bf, c["shapes"], alays)]
    by = [_global_faces(fl, g, s, 1, al.boxes) for fl, g, s, al in zip(flays, bf, c["shapes"], alays)]
    # the fine level's cells and the covered coarse cells
    fine = np.zeros(shp1, dtype=bool)
    for lo, hi in alays[1].boxes:
        fine[lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    cov = fine.reshape(shp0[0], 2, shp0[1], 2).any(axis=(1, 3))
    r0 = alays[0].unpack(rhs[0], shp0)[0]
    r1 = alays[1].unpack(rhs[1], shp1)[0]
    sv = h0 * h0 * r0[~cov].sum() + h1 * h1 * r1[fine].sum()
    mean = sv / (h0 * h0 * (~cov).sum() + h1 * h1 * fine.sum())
    # b on the faces between two fine cells (and two uncovered coarse cells), evaluated here
    A1 = alays[1].unpack(alpha[1], shp1)[0]

    def sharp(a):
        a = np.clip(a, 1e-12, 1 - 1e-12)
        psi = eta / math.sqrt(2.0) * np.log(a / (1 - a))
        return 0.5 + 0.5 * np.tanh(psi / (math.sqrt(2.0) * 2.0 * eta))
    S1 = sharp(A1)
    fx = 0.5 * S1[:, :-1] + 0.5 * S1[:, 1:]
    inner = fine[:, :-1] & fine[:, 1:]
    assert np.max(np.abs((fx * (1 - fx) + 1e-3)[inner] - bx[1][:, 1:-1][inner])) < 1e-14
    worst = 0.0
    # fine composite cells: sum of fine face fluxes
    for j in range(shp1[0]):
        for i in range(shp1[1]):
            if not fine[j, i]:
                continue
            fl = (bx[1][j, i + 1] * gx[1][j, i + 1] - bx[1][j, i] * gx[1][j, i]) * h1 + \
                 (by[1][j + 1, i] * gy[1][j + 1, i] - by[1][j, i] * gy[1][j, i]) * h1
            worst = max(worst, abs(fl - h1 * h1 * (r1[j, i] - mean)))
    # coarse composite cells: a face next to a covered cell takes the two fine faces on it
    def cflux_x(j, i):   # flux through coarse x-face i of row j (positive +x)
        ci, cm = i, i - 1
        if (0 <= ci < shp0[1] and cov[j, ci]) or (0 <= cm < shp0[1] and cov[j, cm]):
            return sum(bx[1][2 * j + o, 2 * i] * gx[1][2 * j + o, 2 * i] for o in range(2)) * h1
        return bx[0][j, i] * gx[0][j, i] * h0

    def cflux_y(j, i):
        if (0 <= j < shp0[0] and cov[j, i]) or (0 <= j - 1 < shp0[0] and cov[j - 1, i]):
            return sum(by[1][2 * j, 2 * i + o] * gy[1][2 * j, 2 * i + o] for o in range(2)) * h1
        return by[0][j, i] * gy[0][j, i] * h0
    for j in range(shp0[0]):
        for i in range(shp0[1]):
            if cov[j, i]:
                continue
            fl = cflux_x(j, i + 1) - cflux_x(j, i) + cflux_y(j + 1, i) - cflux_y(j, i)
            worst = max(worst, abs(fl - h0 * h0 * (r0[j, i] - mean)))
    assert worst <= 1e-9 * h0 * h0 * np.max(np.abs(r0)), worst
    # Neumann domain faces
    for l in range(2):
        assert np.all(np.nan_to_num(gx[l][:, 0]) == 0.0) and np.all(np.nan_to_num(gx[l][:, -1]) == 0.0)
    # a coarse face under the fine level holds the mean of the fine faces on it
    for j in range(shp0[0]):
        for i in range(1, shp0[1]):
            if cov[j, i] and cov[j, i - 1]:
                assert abs(gx[0][j, i] - 0.5 * (gx[1][2 * j, 2 * i] + gx[1][2 * j + 1, 2 * i])) <= 1e-14 * (
                    1.0 + abs(gx[0][j, i]))


def test_advance_with_the_phase_field_step_on_a_hierarchy():
    """Flow + Phase-Field steps of a two-level hierarchy: the flow part's dt history is the flow-only
    hierarchy's for the first step, and totals stay conserved."""
    c = cases.interface_hierarchy(2, 16, [CF_CROSSING], width=2.0)
    P, G, lays, bufs = _hier(c)
    pf = make_pf_params(mechanism="CAC-Adv", explicit_rc=0, eta_multiplier=0.5, cfl=0.3)
    flow = [b.copy() for b in bufs]
    h_flow = oracle.amr_advance(P, G, lays, flow, 1, 0.4, 2)
    tot0 = lays[0].unpack(bufs[0], c["shapes"][0]).reshape(P.nstate, -1).sum(axis=1)
    h = oracle.amr_advance_pf(P, pf, G, lays, bufs, 3, 0.4, 2)
    assert h[0] == h_flow[0] and np.all(h > 0.0)
    U0 = lays[0].unpack(bufs[0], c["shapes"][0])
    tot = U0.reshape(P.nstate, -1).sum(axis=1)
    mag = np.abs(U0).reshape(P.nstate, -1).sum(axis=1)
    assert np.all(np.abs(tot - tot0) <= 1e-13 * mag)
