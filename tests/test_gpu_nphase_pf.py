"""GPU parity of the N-phase model's Phase-Field path (nphase = 2 or 3) through the C ABI against
the oracle on the same inputs: the Phase-Field right-hand side with a Grad_Q component per phase
(mixture and phasic forms, phasic energies clipped at p_min), Grad_Q of every phase (ACDI with the
explicit RC, ACDI and CAC-Adv with the implicit one), the whole Phase-Field step and flow steps
followed by it.

Bars as for the two-phase Phase-Field path: per-variable max relative difference <= 1e-12 for one
evaluation or one step (Grad_Q: max |gpu - cpu| <= 1e-12 max |cpu|; 1e-9 through the implicit
solve, whose sums run in another order), <= 1e-9 in L1 after flow steps; equal sub-step counts;
totals conserved and the volume fractions summing to 1.
"""
import numpy as np
import pytest

import oracle
from compas_amd import FaceLayout, Layout, Level, cases, make_pf_params, scheme_ghosts
from tests._util import l1rel_per_var, maxrel_per_var
from tests.test_oracle_nphase_pf import _rings
from tests.test_oracle_pf import _periodic_random_faces

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12


def _dev(torch, a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def _phase_faces(flay, shape, rng, zero_third=True):
    """Random periodic Grad_Q per phase; the faces of the first third of the domain along x hold
    zero in every phase (faces and whole tile directions without Grad_Q are skipped)."""
    dim = flay.dim
    buf = np.zeros(flay.total)
    for k in range(flay.ncomp):
        faces = _periodic_random_faces(shape, dim, rng, 2.0)
        if zero_third:
            for d in range(dim):
                faces[d][..., : shape[-1] // 3] = 0.0
        for b, (lo, hi) in enumerate(flay.boxes):
            for d in range(dim):
                sl = tuple(slice(lo[e], hi[e] + 1 + (e == d)) for e in reversed(range(dim)))
                flay.face_view(buf, b, d)[k] = faces[d][sl]
    return buf


# (dim, nphase, scheme, form, clip): clip sets p_min of phase 1 inside the pressure range
NP_PF_RHS_CASES = [
    (2, 3, "WENO5", 0, False), (2, 3, "WENO5", 1, False), (2, 3, "WENO5", 1, True), (2, 3, "MUSCL", 1, False),
    (2, 3, "GODUNOV", 0, False), (2, 2, "WENO5", 1, True), (2, 2, "MUSCL", 0, False),
    (3, 3, "WENO5", 1, True), (3, 3, "MUSCL", 0, False), (3, 2, "WENO5", 0, False),
]


@pytest.mark.parametrize("dim,nphase,scheme,form,clip", NP_PF_RHS_CASES)
def test_nphase_pf_rhs_matches_oracle(gpu, dim, nphase, scheme, form, clip):
    import torch
    n = 24 if dim == 2 else 12
    c = cases.nphase_smooth(dim, n, seed=31 + dim + nphase, nphase=nphase, scheme=scheme,
                            max_grid_size=n // 2)
    P, G = c["params"], c["geom"]
    p_min = (1.0, -10.0, -10.0)[:nphase] if clip else (-10.0,) * nphase
    pf = make_pf_params(flux_phase2flux=form, p_min=p_min)
    lay = Layout(dim, c["boxes"], scheme_ghosts(scheme), P.nstate)
    flay = FaceLayout(dim, c["boxes"], nphase)
    buf = lay.pack(c["U"])
    oracle.level_fill_ghosts(P, G, lay, buf)
    gq = _phase_faces(flay, c["shape"], np.random.default_rng(9))
    ref = lay.unpack(oracle.level_pf_rhs(P, pf, G, lay, buf, gq), c["shape"])
    lev = Level(P, G, lay)
    lev.pf_setup(pf)
    assert lev.faces.total == flay.total
    U = _dev(torch, buf, gpu)
    dU = torch.zeros_like(U)
    lev.PhaseField_compute_dUdt_FV(dU, U, _dev(torch, gq, gpu))
    assert lev.check_errors() == 0
    got = lay.unpack(dU.cpu().numpy(), c["shape"])
    lev.close()
    rel = maxrel_per_var(got, ref)
    assert rel.max() <= TOL_STEP, rel
    # the zeroed third (away from its edges) has no Phase-Field flux at all
    inner = got[..., 1: n // 3 - 1]
    assert np.all(inner == 0.0)


# (mechanism, explicit_rc, dim, periodic)
NP_GRADQ_CASES = [("ACDI", 1, 2, True), ("ACDI", 1, 2, False), ("ACDI", 1, 3, True),
                  ("CAC-Adv", 0, 2, True), ("ACDI", 0, 2, False), ("CAC-Adv", 0, 3, False)]


@pytest.mark.parametrize("mech,rc,dim,periodic", NP_GRADQ_CASES)
def test_nphase_pf_gradq_matches_oracle(gpu, mech, rc, dim, periodic):
    import torch
    n = 32 if dim == 2 else 16
    c = _rings(n, dim=dim, width=2.0, periodic=periodic, mg=n // 2)
    P, G = c["params"], c["geom"]
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, eta_multiplier=0.5)
    lay = Layout(dim, c["boxes"], 3, P.nstate)
    flay = FaceLayout(dim, c["boxes"], 3)
    buf = lay.pack(c["U"])
    dt = 0.3 / n
    ref, ref_gm, ref_info = oracle.level_pf_gradq_step(P, pf, G, lay, buf, dt, flay.total)
    lev = Level(P, G, lay)
    lev.pf_setup(pf)
    U = _dev(torch, buf, gpu)
    gq = torch.full((flay.total,), float("nan"), dtype=torch.float64, device=gpu)
    gm, info = lev.PhaseField_GradQPF(U, dt, gq)
    assert lev.check_errors() == 0
    got = gq.cpu().numpy()
    lev.close()
    assert info[3] == ref_info[3] and info[3] >= 1
    for q in range(3):
        assert abs(info[q] - ref_info[q]) <= 1e-15 * abs(ref_info[q])
    assert np.all(np.isfinite(got))
    tol = TOL_STEP if rc == 1 else 1e-9
    for k in range(3):
        r = np.concatenate([flay.face_view(ref, b, d)[k].ravel() for b in range(len(c["boxes"]))
                            for d in range(dim)])
        g = np.concatenate([flay.face_view(got, b, d)[k].ravel() for b in range(len(c["boxes"]))
                            for d in range(dim)])
        scale = np.max(np.abs(r))
        assert scale > 0.0, k
        assert np.max(np.abs(g - r)) <= tol * scale, (k, np.max(np.abs(g - r)) / scale)
    assert abs(gm - ref_gm) <= tol * ref_gm


@pytest.mark.parametrize("mech,rc,dim,form", [("ACDI", 1, 2, 0), ("ACDI", 1, 2, 1), ("CAC-Adv", 0, 2, 0),
                                              ("ACDI", 1, 3, 0)])
def test_nphase_pf_evolve_matches_oracle(gpu, mech, rc, dim, form):
    import torch
    n = 32 if dim == 2 else 16
    c = _rings(n, dim=dim, width=2.5, mg=n // 2)
    P, G = c["params"], c["geom"]
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, flux_phase2flux=form, eta_multiplier=0.5, cfl=0.3,
                        p_min=(-10.0, -10.0, -10.0))
    lay = Layout(dim, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    dt = 2.0 / n
    Un = _dev(torch, buf, gpu)
    ref_info = oracle.level_pf_evolve(P, pf, G, lay, buf, dt)
    ref = lay.unpack(buf, c["shape"])
    lev = Level(P, G, lay)
    lev.pf_setup(pf)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dtd = torch.full((1,), dt, dtype=torch.float64, device=gpu)
    info = lev.PhaseField_Evolve(Un, U1, U2, dtd)
    assert lev.check_errors() == 0
    got = lay.unpack(Un.cpu().numpy(), c["shape"])
    lev.close()
    assert info[3] == ref_info[3] and info[5] == ref_info[5] and info[5] >= 2
    rel = maxrel_per_var(got, ref)
    assert rel.max() <= (TOL_STEP if rc == 1 else 1e-10), rel
    for q in range(P.nstate):
        assert abs(got[q].sum() - c["U"][q].sum()) <= 1e-12 * np.abs(c["U"][q]).sum()
    ia = 4 + dim
    assert np.max(np.abs(got[ia:ia + 3].sum(axis=0) - 1.0)) <= 1e-12


def test_nphase_advance_with_pf_matches_oracle(gpu):
    """Ten flow steps of the three-phase rings (48^2 on a 2x2 split, uniform flow), each followed by
    the CAC-Adv Phase-Field step with the implicit RC, graph-replayed flow part: L1 <= 1e-9, equal
    dt history, totals conserved."""
    import torch
    n = 48
    c = _rings(n, width=2.0, mg=n // 2)
    P, G = c["params"], c["geom"]
    pf = make_pf_params(mechanism="CAC-Adv", explicit_rc=0, eta_multiplier=0.5, cfl=0.3)
    lay = Layout(2, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    hist_ref = oracle.level_advance_pf(P, pf, G, lay, buf, 10, 0.4)
    ref = lay.unpack(buf, c["shape"])
    lev = Level(P, G, lay)
    lev.pf_setup(pf)
    Un = _dev(torch, lay.pack(c["U"]), gpu)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dt = torch.zeros(1, dtype=torch.float64, device=gpu)
    cm = torch.zeros(1, dtype=torch.float64, device=gpu)
    hist = torch.zeros(10, dtype=torch.float64, device=gpu)
    lev.advance(Un, U1, U2, 10, 0.4, dt, cm, hist, use_graph=True)
    assert lev.check_errors() == 0
    got = lay.unpack(Un.cpu().numpy(), c["shape"])
    lev.close()
    assert np.max(np.abs(hist.cpu().numpy() - hist_ref) / hist_ref) <= 1e-12
    l1 = l1rel_per_var(got, ref)
    assert l1.max() <= 1e-9, l1
    for q in range(P.nstate):
        assert abs(got[q].sum() - c["U"][q].sum()) <= 1e-12 * np.abs(c["U"][q]).sum()
    assert np.max(np.abs(got[6:9].sum(axis=0) - 1.0)) <= 1e-12


@pytest.mark.parametrize("dim,nphase,scheme,form", [(2, 0, "WENO5", 1), (2, 0, "WENO5B", 0), (3, 0, "WENO5", 1),
                                                    (2, 3, "WENO5", 0), (2, 3, "MUSCL", 1), (3, 3, "WENO5", 1)])
def test_pf_rhs_box_equals_level_form(gpu, dim, nphase, scheme, form):
    """cmp_pf_rhs_box (one box, its fab's ghosts filled, its own Grad_Q face block) gives every box
    the same right-hand side bit for bit as cmp_level_pf_rhs over the level, which the oracle pins;
    the box with a zero Grad_Q third gets exact zeros there."""
    import torch
    from compas_amd import pf_rhs_box
    n = 24 if dim == 2 else 12
    if nphase:
        c = cases.nphase_smooth(dim, n, seed=41 + dim, nphase=nphase, scheme=scheme, max_grid_size=n // 2)
        pf = make_pf_params(flux_phase2flux=form, p_min=(-10.0,) * nphase)
    else:
        c = cases.random_smooth(dim, n, seed=43 + dim, scheme=scheme, max_grid_size=n // 2)
        pf = make_pf_params(flux_phase2flux=form, p_min=(-1.0e3, -1.0e3))
    P, G = c["params"], c["geom"]
    ng = scheme_ghosts(scheme)
    lay = Layout(dim, c["boxes"], ng, P.nstate)
    flay = FaceLayout(dim, c["boxes"], nphase or 1)
    buf = lay.pack(c["U"])
    oracle.level_fill_ghosts(P, G, lay, buf)
    gq = _phase_faces(flay, c["shape"], np.random.default_rng(17))
    lev = Level(P, G, lay)
    lev.pf_setup(pf)
    dU = torch.zeros(lay.total, dtype=torch.float64, device=gpu)
    lev.PhaseField_compute_dUdt_FV(dU, _dev(torch, buf, gpu), _dev(torch, gq, gpu))
    assert lev.check_errors() == 0
    lev.close()
    lvl = dU.cpu().numpy()
    for b, (lo, hi) in enumerate(c["boxes"]):
        fab = lay.fab_view(buf, b)
        g0 = flay.offsets[b][0]
        g1 = flay.offsets[b + 1][0] if b + 1 < len(c["boxes"]) else flay.total
        out = torch.full((P.nstate * int(np.prod([hi[d] - lo[d] + 1 for d in range(dim)])),), float("nan"),
                         dtype=torch.float64, device=gpu)
        pf_rhs_box(P, pf, G, (lo, hi), ng, _dev(torch, fab.ravel(), gpu), _dev(torch, gq[g0:g1], gpu), out)
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape((P.nstate,) + tuple(hi[d] - lo[d] + 1 for d in reversed(range(dim))))
        sl = (slice(None),) + tuple(slice(ng, ng + hi[d] - lo[d] + 1) for d in reversed(range(dim)))
        want = lay.fab_view(lvl, b)[sl]
        assert np.array_equal(got, want), (b, np.max(np.abs(got - want)))
