"""Pins of the oracle's N-phase Phase-Field step (CPU only):

- nphase = 2 with Grad_Q = (g, -g) is the two-phase model: the efficient UD flux's Sum_J vanishes
  exactly and the flux equals the two-phase UD flux (mixture and phasic forms, with and without
  phasic energies clipped at p_min);
- the phasic split of one state conserves mass, momentum and energy; a clipped phase sits at its
  p_min energy and the unclipped ones share the remainder in proportion to their volume fractions;
- the volume-fraction fluxes sum to zero (Sum_J is taken back out upwind), so sum(alpha_k) stays 1;
- a whole N-phase Phase-Field step (CAC-Adv and ACDI, implicit and explicit RC) on a three-phase
  ring conserves every total, keeps uniform pressure and velocity uniform and sum(alpha) = 1;
- the nphase = 2 level step equals the two-phase level step on the same flow.
"""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, cases, make_geometry, make_params, make_pf_params, split_domain
from tests.test_oracle_pf import _interface_case


def _two_phase_state(rng, P, dim, pmul=1.0):
    a1 = rng.uniform(1e-6, 1 - 1e-6)
    W = np.empty(4 + dim)
    W[0], W[1] = a1 * rng.uniform(0.5, 2.0), (1 - a1) * rng.uniform(5.0, 12.0)
    W[2:2 + dim] = rng.uniform(-1, 1, dim)
    W[2 + dim] = pmul * rng.uniform(0.5, 2.0)
    W[3 + dim] = a1
    return cases.prim_to_cons(P, W.reshape(-1, 1)).ravel()


def _to_np2(P2, U):
    dim = P2.dim
    out = np.empty(5 + dim)
    out[:4 + dim] = U
    out[4 + dim] = 1.0 - U[3 + dim]
    return out


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("phase2flux", [0, 1])
@pytest.mark.parametrize("clip", [False, True])
def test_np2_flux_is_the_two_phase_flux(dim, phase2flux, clip):
    rng = np.random.default_rng(10 * dim + phase2flux + 100 * clip)
    P2 = make_params(dim, gamma=(1.4, 1.6), pinf=(0.0, 0.0))
    Pn = make_params(dim, gamma=(1.4, 1.6), pinf=(0.0, 0.0), nphase=2)
    # p_min above the pressure forces the clipping of one phase (or both: then the loop stops)
    p_min = (1.5, -100.0) if clip else (0.0, 0.0)
    pf = make_pf_params(flux_phase2flux=phase2flux, p_min=p_min)
    for _ in range(40):
        UL, UR = _two_phase_state(rng, P2, dim), _two_phase_state(rng, P2, dim)
        g = rng.uniform(-2, 2)
        F2, c2 = oracle.pf_flux(P2, pf, UL, UR, g)
        Fn, cn = oracle.np_pf_flux(Pn, pf, _to_np2(P2, UL), _to_np2(P2, UR), [g, -g])
        # two-phase order: a1r1, a2r2, mom, E, a1; N-phase: a1r1, a2r2, mom, E, a1, a2
        ref = np.concatenate([F2, [-F2[3 + dim]]])
        assert np.all(np.abs(Fn - ref) <= 1e-12 * np.abs(ref) + 1e-14 * np.max(np.abs(ref))), (Fn, ref)
        assert abs(cn - c2) <= 1e-14 * c2


def test_np_cons2phase_conserves_and_clips():
    dim = 2
    P = make_params(dim, gamma=(1.4, 1.67, 1.09), pinf=(0.0, 0.5, 0.0), nphase=3)
    rng = np.random.default_rng(3)
    nclipped = 0
    for t in range(60):
        a = rng.dirichlet([1.0, 1.0, 1.0])
        W = np.empty(2 * 3 + dim + 1)
        W[:3] = a * rng.uniform(0.5, 3.0, 3)
        W[3:3 + dim] = rng.uniform(-1, 1, dim)
        W[3 + dim] = rng.uniform(0.5, 2.0)
        W[4 + dim:] = a
        U = cases.np_prim_to_cons(P, W.reshape(-1, 1)).ravel()
        p = W[3 + dim]
        p_min = tuple(rng.choice([-10.0, 0.9 * p, 1.2 * p, 3.0 * p]) for _ in range(3))
        pf = make_pf_params(flux_phase2flux=0, p_min=p_min)
        Cp = oracle.np_cons2phase(P, pf, U)
        mass, E, alpha = Cp[0:3], Cp[(1 + dim) * 3:(2 + dim) * 3], Cp[(2 + dim) * 3:]
        assert np.array_equal(mass, U[:3]) and np.array_equal(alpha, U[4 + dim:])
        for m in range(dim):
            assert abs(Cp[(1 + m) * 3:(2 + m) * 3].sum() - U[3 + m]) <= 1e-14 * (abs(U[3 + m]) + 1)
        ke = 0.5 * W[:3] * sum(W[3 + m] ** 2 for m in range(dim))
        e = E - ke
        A, B = np.array(P.eos_A[:3]), np.array(P.eos_B[:3])
        e_min = a * (A * np.array(p_min) + B)
        e_eq = a * (A * p + B)
        clipped = e_eq < e_min
        at_min = np.abs(e - e_min) <= 1e-13 * np.abs(e_min) + 1e-15
        if at_min.all():
            continue   # every phase clipped: no energy is left to share (the reference does the same)
        assert abs(E.sum() - U[3 + dim]) <= 1e-13 * abs(U[3 + dim])
        if clipped.any() and not clipped.all():
            nclipped += 1
            # clipped phases first sit at e_min; the loop can clip more, never fewer
            assert np.all(np.abs(e[clipped] - e_min[clipped]) <= 1e-13 * np.abs(e_min[clipped]) + 1e-15)
            free = e > e_min * (1 + 1e-12) + 1e-15
            if free.sum() >= 2:
                ratio = e[free] - e_eq[free]
                wts = a[free] / a[free].sum()
                share = ratio / ratio.sum()
                assert np.allclose(share, wts, rtol=1e-10)
        elif not clipped.any():
            assert np.allclose(e, e_eq, rtol=1e-13)
    assert nclipped >= 10


@pytest.mark.parametrize("phase2flux", [0, 1])
def test_np_alpha_fluxes_sum_to_zero(phase2flux):
    dim = 3
    P = make_params(dim, gamma=(1.4, 1.67, 1.09), pinf=(0.0, 0.0, 0.0), nphase=3)
    pf = make_pf_params(flux_phase2flux=phase2flux)
    rng = np.random.default_rng(7)
    for _ in range(40):
        Us = []
        for s in range(2):
            a = rng.dirichlet([2.0, 2.0, 2.0])
            W = np.concatenate([a * rng.uniform(0.5, 3.0, 3), rng.uniform(-1, 1, dim),
                                [rng.uniform(0.5, 2.0)], a])
            Us.append(cases.np_prim_to_cons(P, W.reshape(-1, 1)).ravel())
        g = rng.uniform(-1, 1, 3)
        F, c = oracle.np_pf_flux(P, pf, Us[0], Us[1], g)
        fa = F[4 + dim:]
        assert abs(fa.sum()) <= 1e-15 * np.abs(fa).sum() * 8 + 1e-300
        assert c > 0


def _rings(n, dim=2, width=1.5, periodic=True, mg=None):
    """Three ideal-gas phases: a disk (phase 1) inside a ring (phase 3) inside phase 2, uniform u, p."""
    P = make_params(dim, gamma=(1.4, 1.67, 1.09), pinf=(0.0, 0.0, 0.0), nphase=3)
    shape = (n,) * dim
    G = make_geometry(shape, (0.0,) * dim, (1.0,) * dim, (int(periodic),) * dim)
    X = cases._centres(G, shape)
    r = np.sqrt(sum((X[d] - 0.5) ** 2 for d in range(dim)))
    h1 = 0.5 * (1 + np.tanh((0.2 - r) / (width * G.dx[0])))
    h2 = 0.5 * (1 + np.tanh((0.32 - r) / (width * G.dx[0])))
    a = [np.clip(v, 1e-8, 1.0) for v in (h1, 1.0 - h2, h2 - h1)]
    tot = sum(a)
    a = [v / tot for v in a]
    W = np.empty((P.nstate,) + shape)
    for k, rho in enumerate((1.0, 0.14, 5.0)):
        W[k] = a[k] * rho
        W[4 + dim + k] = a[k]
    for m in range(dim):
        W[3 + m] = (1.0, 0.5, -0.25)[m]
    W[3 + dim] = 1.0
    U = cases.np_prim_to_cons(P, W)
    return dict(params=P, geom=G, U=U, boxes=split_domain(shape, mg or n), shape=shape)


def _np_prims(P, U):
    dim, np_ = P.dim, P.nphase
    rho = sum(U[k] for k in range(np_))
    vel = [U[np_ + m] / rho for m in range(dim)]
    rho_e = U[np_ + dim] - 0.5 * rho * sum(v * v for v in vel)
    AP = sum(U[np_ + dim + 1 + k] * P.eos_A[k] for k in range(np_))
    BP = sum(U[np_ + dim + 1 + k] * P.eos_B[k] for k in range(np_))
    return vel, (rho_e - BP) / AP


@pytest.mark.parametrize("mechanism,explicit_rc", [("CAC-Adv", 0), ("ACDI", 1), ("ACDI", 0)])
def test_np_pf_step_conserves_and_keeps_equilibrium(mechanism, explicit_rc):
    c = _rings(24, mg=12)
    P, G = c["params"], c["geom"]
    lay = Layout(2, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    pf = make_pf_params(mechanism=mechanism, explicit_rc=explicit_rc, eta_multiplier=0.5, cfl=0.3)
    info = oracle.level_pf_evolve(P, pf, G, lay, buf, 0.004)
    assert info[5] >= 1 and info[4] > 0.0
    U = lay.unpack(buf, c["shape"])
    for q in range(P.nstate):
        assert abs(U[q].sum() - c["U"][q].sum()) <= 1e-12 * np.abs(c["U"][q]).sum()
    vel, p = _np_prims(P, U)
    assert np.max(np.abs(vel[0] - 1.0)) < 1e-11 and np.max(np.abs(vel[1] - 0.5)) < 1e-11
    assert np.max(np.abs(p - 1.0)) < 1e-11
    assert np.max(np.abs(U[6:9].sum(axis=0) - 1.0)) < 1e-13
    for k in range(3):
        assert np.max(np.abs(U[6 + k] - c["U"][6 + k])) > 1e-7


@pytest.mark.parametrize("mechanism,explicit_rc", [("ACDI", 1), ("CAC-Adv", 0), ("ACDI", 0)])
def test_np2_pf_step_is_the_two_phase_step(mechanism, explicit_rc):
    """Phase 2's mechanism and Grad_Q, computed from alpha_2 = 1 - alpha_1, are the negated phase-1
    ones to round-off, so the two models agree to a few ulps (1e-13 leaves room for the solver)."""
    tol = 1e-13
    c2 = _interface_case(24, width=2.0, max_grid_size=12)
    cn = cases.nphase_from_two_phase(c2, nphase=2)
    pf = make_pf_params(mechanism=mechanism, explicit_rc=explicit_rc, eta_multiplier=0.5, cfl=0.3)
    out = []
    for c in (c2, cn):
        P = c["params"]
        lay = Layout(2, c["boxes"], 3, P.nstate)
        buf = lay.pack(c["U"])
        info = oracle.level_pf_evolve(P, pf, c["geom"], lay, buf, 0.004)
        out.append((lay.unpack(buf, c["shape"]), info))
    (U2, i2), (Un, iN) = out
    assert i2[5] == iN[5] and abs(i2[4] - iN[4]) <= 1e-9 * i2[4]
    pairs = [(0, 0), (1, 1), (2, 2), (3, 3), (4, 4), (5, 5)]
    for a, b in pairs:
        ref = U2[a]
        err = np.max(np.abs(Un[b] - ref)) / np.max(np.abs(ref))
        assert err <= tol, (a, err)
    assert np.max(np.abs(Un[6] - (1.0 - U2[5]))) <= tol
