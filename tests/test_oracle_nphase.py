"""Pins of the oracle's N-phase model (nphase = 2 or 3; CPU only), with checks that do not restate
its formulas:

- a three-phase state whose phases 2 and 3 are the same material (same EOS, same density) is the
  two-phase model's state with those phases merged: mixture EOS, sound speed (both K modes), the
  physical flux, HLLC/HLL fluxes and u*, and the Godunov RHS of a level all agree with the two-phase
  oracle after merging;
- nphase = 2 with WENO5: the stored alpha2 is 1 - alpha1, the alpha slopes cancel, the consistent
  limiter leaves every face alone, and the whole RHS equals the two-phase model's;
- Wood's mixture: sum_k alpha_k (1 + K_k) = 1;
- the consistent limiter: sum_k theta_k s_k = 0 with theta in [0, 1] when it acts, theta = 1 below
  its threshold or when the slopes already sum to zero, the first phase of largest |slope| on ties;
  a limited face has volume fractions summing to 1 when the cell's do;
- a periodic level conserves the phase masses, momentum and energy, and sum_k d(alpha_k)/dt = 0
  (K = 0), which needs the limiter;
- uniform pressure and velocity stay uniform across the interfaces through an RK3 step;
- reflecting ghost cells flip the momentum components, which sit after the phase masses.
"""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, cases, make_geometry, make_params, split_domain


def _merge3(P3, U):
    """three-phase conservative state (phases 2, 3 the same material) -> two-phase state"""
    dim = P3.dim
    out = np.empty((4 + dim,) + U.shape[1:])
    out[0] = U[0]
    out[1] = U[1] + U[2]
    for m in range(dim):
        out[2 + m] = U[3 + m]
    out[2 + dim] = U[3 + dim]
    out[3 + dim] = U[4 + dim]
    return out


def _merge3_rhs(P3, R):
    """merged rates: alpha1 rate from alpha_1's, phase-2 mass rate from both"""
    dim = P3.dim
    out = np.empty((4 + dim,) + R.shape[1:])
    out[0] = R[0]
    out[1] = R[1] + R[2]
    for m in range(dim):
        out[2 + m] = R[3 + m]
    out[2 + dim] = R[3 + dim]
    out[3 + dim] = R[4 + dim]
    return out


def _state3(rng, dim, P3):
    a = rng.uniform(0.1, 0.8)
    a2 = (1.0 - a) * rng.uniform(0.2, 0.8)
    a3 = 1.0 - a - a2
    rho1, rho2 = rng.uniform(0.5, 3.0), rng.uniform(0.5, 3.0)
    W = np.zeros(P3.nstate)
    W[0], W[1], W[2] = a * rho1, a2 * rho2, a3 * rho2
    W[3:3 + dim] = rng.uniform(-1, 1, dim)
    W[3 + dim] = rng.uniform(0.5, 2.0)
    W[4 + dim:] = [a, a2, a3]
    return oracle.np_prim2cons(P3, W)


def _params_pair(dim, riemann="HLLC", source_term=0, pinf=(0.0, 3.0)):
    P2 = make_params(dim, gamma=(1.4, 1.9), pinf=pinf, riemann=riemann, source_term=source_term)
    P3 = make_params(dim, gamma=(1.4, 1.9, 1.9), pinf=(pinf[0], pinf[1], pinf[1]), riemann=riemann,
                     source_term=source_term, nphase=3)
    return P2, P3


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("id_k", [0, 1])
def test_identical_phases_merge_into_the_two_phase_model(dim, id_k):
    rng = np.random.default_rng(5 + dim + 10 * id_k)
    P2, P3 = _params_pair(dim, source_term=id_k)
    for _ in range(20):
        U3 = _state3(rng, dim, P3)
        U2 = _merge3(P3, U3[:, None])[:, 0]
        W3 = oracle.np_cons2prim(P3, U3)
        W2 = oracle.cons2prim(P2, U2)
        assert abs(W3[3 + dim] - W2[2 + dim]) <= 1e-13 * abs(W2[2 + dim])          # pressure
        assert np.allclose(W3[3:3 + dim], W2[2:2 + dim], rtol=1e-14, atol=1e-15)   # velocity
        assert np.allclose(oracle.np_prim2cons(P3, W3), U3, rtol=1e-14, atol=1e-15)
        c3 = oracle.np_sound_speed(P3, W3[3 + dim], W3[4 + dim:], W3[:3], id_k)
        c2 = oracle.sound_speed(P2, W2[2 + dim], W2[3 + dim], W2[0], W2[1], id_k)
        assert abs(c3 - c2) <= 1e-13 * c2
        for d in range(dim):
            F3 = oracle.np_cons2flux(P3, U3, d)
            F2 = oracle.cons2flux(P2, U2, d)
            assert np.allclose(_merge3_rhs(P3, F3[:, None])[:, 0], F2, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("riemann", ["HLLC", "HLL"])
@pytest.mark.parametrize("dim", [2, 3])
def test_identical_phases_riemann(riemann, dim):
    rng = np.random.default_rng(11 + dim)
    P2, P3 = _params_pair(dim, riemann=riemann)
    for _ in range(20):
        UL, UR = _state3(rng, dim, P3), _state3(rng, dim, P3)
        for d in range(dim):
            F3, v3, c3 = oracle.riemann(P3, UL, UR, d)
            F2, v2, c2 = oracle.riemann(P2, _merge3(P3, UL[:, None])[:, 0], _merge3(P3, UR[:, None])[:, 0], d)
            assert np.allclose(_merge3_rhs(P3, F3[:, None])[:, 0], F2, rtol=1e-12, atol=1e-12)
            assert np.allclose(v3, v2, rtol=1e-12, atol=1e-13)
            assert abs(c3 - c2) <= 1e-13 * c2
            # consistency: F(U, U) is the physical flux, u* the velocity
            F, v, _ = oracle.riemann(P3, UL, UL, d)
            assert np.allclose(F, oracle.np_cons2flux(P3, UL, d), rtol=1e-13, atol=1e-13)


def test_wood_mixture_identity():
    """K = 1: sum_k alpha_k (1 + K_k) = rho c^2 sum_k alpha_k/(rho_k c_k^2) = 1 (no covolume)."""
    rng = np.random.default_rng(2)
    P = make_params(2, gamma=(1.4, 1.67, 4.4), pinf=(0.0, 0.0, 6.0), source_term=1, nphase=3)
    for _ in range(50):
        alpha = rng.dirichlet([1.0, 1.0, 1.0])
        arho = alpha * rng.uniform(0.5, 5.0, 3)
        p = rng.uniform(0.5, 5.0)
        K = oracle.np_K(P, p, alpha, arho, 1)
        assert abs(np.sum(alpha * (1.0 + K)) - 1.0) <= 1e-14
        assert np.all(oracle.np_K(P, p, alpha, arho, 0) == 0.0)
        # a pure phase: the textbook gamma (p + pinf)/rho
        c = oracle.np_sound_speed(P, p, [0.0, 0.0, 1.0], [0.0, 0.0, 1000.0], 0)
        assert abs(c * c - 4.4 * (p + 6.0) / 1000.0) <= 1e-14 * c * c


def test_consistent_limiter_closes_the_slopes():
    rng = np.random.default_rng(7)
    for _ in range(500):
        s = rng.uniform(-1, 1, 3)
        th = oracle.np_limiter_consistent(s)
        assert np.all((th >= 0.0) & (th <= 1.0))
        assert abs(np.dot(th, s)) <= 1e-14 * np.max(np.abs(s))
    # slopes that already close, and slopes below the threshold: no limiting
    assert np.all(oracle.np_limiter_consistent([0.3, -0.1, -0.2]) == 1.0)
    assert np.all(oracle.np_limiter_consistent([1e-11, 2e-11, -5e-12]) == 1.0)
    # a tie in |slope| picks the first phase: (s_max - S)/s_max for it, s_max/(s_max - S) (clipped)
    th = oracle.np_limiter_consistent([0.5, -0.5, 0.25])
    assert th[0] == pytest.approx(0.5) and th[1] == 1.0 and th[2] == 1.0
    # two phases with s2 = -s1 + e: one of them is scaled to close the sum
    th = oracle.np_limiter_consistent([0.4, -0.3])
    assert abs(0.4 * th[0] - 0.3 * th[1]) <= 1e-15


def test_limited_face_fractions_sum_to_one():
    rng = np.random.default_rng(9)
    P = make_params(2, gamma=(1.4, 1.67, 1.09), nphase=3)
    for _ in range(50):
        Uc = _state3(rng, 2, P)
        Wc = oracle.np_cons2prim(P, Uc)
        Wf = Wc.copy()
        # primitives: alpha_k rho_k 0..2, u v 3..4, p 5, alpha_k 6..8; an inconsistent face
        Wf[6:] = np.clip(Wc[6:] + rng.uniform(-0.05, 0.05, 3), 0.01, 0.99)
        Wf[:3] = Wf[6:] * Wc[:3] / Wc[6:]
        Uf = oracle.np_prim2cons(P, Wf)
        for face in (1, -1):
            Uo = oracle.np_limit_consistent(P, Uf, Uc, 0.1, face)
            Wo = oracle.np_cons2prim(P, Uo)
            assert abs(np.sum(Wo[6:]) - 1.0) <= 2e-15
            assert np.allclose(Wo[3:6], Wf[3:6], rtol=1e-13, atol=1e-14)   # u, v, p untouched


def _level_rhs(c):
    P, G = c["params"], c["geom"]
    lay = Layout(P.dim, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    oracle.level_fill_ghosts(P, G, lay, buf)
    dU, cmax = oracle.level_rhs(P, G, lay, buf)
    return lay.unpack(dU, c["shape"]), cmax


@pytest.mark.parametrize("dim,riemann", [(2, "HLLC"), (2, "HLL"), (3, "HLLC")])
def test_godunov_three_phase_level_equals_two_phase(dim, riemann):
    n = 16 if dim == 2 else 8
    c2 = cases.random_smooth(dim, n, seed=4, scheme="GODUNOV", riemann=riemann, max_grid_size=n // 2)
    c3 = cases.nphase_from_two_phase(c2, 3)
    r2, cm2 = _level_rhs(c2)
    r3, cm3 = _level_rhs(c3)
    m = _merge3_rhs(c3["params"], r3)
    for q in range(r2.shape[0]):
        assert np.max(np.abs(m[q] - r2[q])) <= 1e-12 * max(np.max(np.abs(r2[q])), 1e-300), q
    assert abs(cm3 - cm2) <= 1e-13 * cm2


@pytest.mark.parametrize("dim,scheme", [(2, "WENO5"), (2, "MUSCL"), (3, "WENO5")])
def test_two_phase_n_phase_model_equals_two_phase_model(dim, scheme):
    """nphase = 2: the limiter is a no-op (alpha2 = 1 - alpha1 everywhere) and the RHS is the
    two-phase model's; the alpha2 rate is minus the alpha1 rate."""
    n = 16 if dim == 2 else 8
    c2 = cases.random_smooth(dim, n, seed=6, scheme=scheme, max_grid_size=n // 2)
    cN = cases.nphase_from_two_phase(c2, 2)
    r2, _ = _level_rhs(c2)
    rN, _ = _level_rhs(cN)
    m = np.empty_like(r2)
    m[0], m[1] = rN[0], rN[1]
    for q in range(dim + 1):
        m[2 + q] = rN[2 + q]
    m[3 + dim] = rN[3 + dim]
    for q in range(r2.shape[0]):
        scale = max(np.max(np.abs(r2[q])), 1e-300)
        assert np.max(np.abs(m[q] - r2[q])) <= 1e-12 * scale, q
    assert np.max(np.abs(rN[4 + dim] + rN[3 + dim])) <= 1e-12 * np.max(np.abs(rN[3 + dim]))


@pytest.mark.parametrize("dim,scheme,riemann", [(2, "WENO5", "HLLC"), (2, "MUSCL", "HLL"),
                                                (3, "WENO5", "HLLC")])
def test_three_phase_level_conserves_and_keeps_fractions_closed(dim, scheme, riemann):
    n = 16 if dim == 2 else 8
    c = cases.nphase_smooth(dim, n, seed=3, scheme=scheme, riemann=riemann, max_grid_size=n // 2)
    P = c["params"]
    r, _ = _level_rhs(c)
    for q in range(P.nphase + dim + 1):
        assert abs(r[q].sum()) <= 1e-12 * np.abs(r[q]).sum(), q
    ia = P.nphase + dim + 1
    sa = r[ia:ia + 3].sum(axis=0)
    assert np.max(np.abs(sa)) <= 1e-12 * np.max(np.abs(r[ia]))
    # without the limiter the fractions' rates do not close: the check above is not vacuous
    assert np.max(np.abs(r[ia])) > 1e-3


def test_three_phase_rk3_step_keeps_pressure_and_velocity_uniform():
    n = 24
    P = make_params(2, gamma=(1.4, 1.67, 1.09), pinf=(0.0, 1.0, 0.0), nphase=3)
    G = make_geometry((n, n), (0.0, 0.0), (1.0, 1.0), (1, 1))
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="xy")
    w1 = 0.5 + 0.5 * np.tanh((0.25 - np.hypot(X - 0.5, Y - 0.5)) / 0.05)
    w2 = 0.5 + 0.5 * np.tanh((X - 0.5) / 0.05)
    a = [np.clip(w1, 1e-6, 1), np.clip((1 - w1) * w2, 1e-6, 1), np.clip((1 - w1) * (1 - w2), 1e-6, 1)]
    tot = a[0] + a[1] + a[2]
    a = [v / tot for v in a]
    W = np.empty((9, n, n))
    for k, rho in enumerate((1.0, 0.14, 5.0)):
        W[k] = a[k] * rho
        W[6 + k] = a[k]
    W[3], W[4], W[5] = 1.0, 0.5, 1.0
    U = cases.np_prim_to_cons(P, W)
    lay = Layout(2, split_domain((n, n), 12), 3, 9)
    buf = lay.pack(U)
    oracle.level_rk3_step(P, G, lay, buf, 0.2 / n)
    V = lay.unpack(buf, (n, n))
    Wn = np.stack([oracle.np_cons2prim(P, V[:, j, i]) for j in range(n) for i in range(n)])
    assert np.max(np.abs(Wn[:, 3] - 1.0)) < 1e-12
    assert np.max(np.abs(Wn[:, 4] - 0.5)) < 1e-12
    assert np.max(np.abs(Wn[:, 5] - 1.0)) < 1e-11
    assert np.max(np.abs(V[6:9] - U[6:9])) > 1e-4


def test_reflecting_fill_flips_the_n_phase_momenta():
    n = 8
    c = cases.nphase_smooth(2, n, seed=1)
    P = c["params"]
    G = make_geometry((n, n), (0.0, 0.0), (1.0, 1.0), (0, 0), lo_bc=(3, 5), hi_bc=(4, 3))
    lay = Layout(2, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    oracle.level_fill_ghosts(P, G, lay, buf)
    fab = buf.reshape(P.nstate, n + 6, n + 6)
    U = c["U"]
    # x-low Symmetry: u odd, v even; x-high SlipWall: the same
    assert np.array_equal(fab[3, 3:-3, 2], -U[3, :, 0]) and np.array_equal(fab[4, 3:-3, 2], U[4, :, 0])
    assert np.array_equal(fab[3, 3:-3, -3], -U[3, :, -1])
    # y-low NoSlipWall: both odd; masses and fractions even
    assert np.array_equal(fab[3, 2, 3:-3], -U[3, 0, :]) and np.array_equal(fab[4, 2, 3:-3], -U[4, 0, :])
    for q in (0, 1, 2, 5, 6, 7, 8):
        assert np.array_equal(fab[q, 2, 3:-3], U[q, 0, :])
