"""Pins the oracle's pointwise functions against facts that do not come from its own formulas:
textbook stiffened-gas thermodynamics, the thermodynamic definition of the sound speed, the
textbook Kapila coefficient, polynomial exactness and convergence order of WENO5, classical TVD
limiter definitions and Toro's HLLC for the single-phase Euler equations. CPU only.

A reference golden vector does not exist for any of these functions (the reference ships no unit
tests or fixtures for the path); each check below is chosen so that a dropped term, a wrong sign,
a swapped index or a transposed operand in the oracle fails it.
"""
import math

import numpy as np
import pytest

import oracle
from compas_amd import make_params


def sg_params(dim=2, gamma=(4.4, 1.4), pinf=(6.0e8, 0.0), source_term=0):
    return make_params(dim, gamma=gamma, pinf=pinf, source_term=source_term)


def state(dim, a1r1, a2r2, vel, p, a1, params):
    W = np.array([a1r1, a2r2] + list(vel) + [p, a1], dtype=np.float64)
    return oracle.prim2cons(params, W), W


# ---- EOS -----------------------------------------------------------------------------------------

def test_pure_phase_pressure_is_stiffened_gas():
    """alpha1 = 1: p = (gamma - 1) rho e - gamma p_inf (textbook stiffened gas)."""
    P = sg_params()
    rho, u, v, rhoe = 998.0, 3.0, -2.0, 2.4e9
    E = rhoe + 0.5 * rho * (u * u + v * v)
    U = np.array([rho, 0.0, rho * u, rho * v, E, 1.0])
    W = oracle.cons2prim(P, U)
    g, pinf = 4.4, 6.0e8
    assert W[4] == pytest.approx((g - 1.0) * rhoe - g * pinf, rel=1e-13)
    assert W[2] == pytest.approx(u, rel=1e-15) and W[3] == pytest.approx(v, rel=1e-15)


def test_mixture_pressure_is_isobaric_closure():
    """Mixture rho e = sum alpha_k (p + gamma_k pinf_k)/(gamma_k - 1): recover p from it."""
    P = sg_params()
    a1, p = 0.37, 2.5e6
    rhoe = a1 * (p + 4.4 * 6.0e8) / 3.4 + (1 - a1) * (p + 0.0) / 0.4
    rho1, rho2 = 1000.0, 1.2
    U = np.array([a1 * rho1, (1 - a1) * rho2, 0.0, 0.0, rhoe, a1])
    assert oracle.cons2prim(P, U)[4] == pytest.approx(p, rel=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_prim_cons_round_trip(dim):
    P = sg_params(dim, pinf=(3.0, 0.0), gamma=(1.6, 1.4))
    rng = np.random.default_rng(1)
    for _ in range(50):
        a1 = rng.uniform(1e-8, 1 - 1e-8)
        W = np.array([a1 * rng.uniform(0.5, 2), (1 - a1) * rng.uniform(0.5, 2)]
                     + list(rng.uniform(-2, 2, size=dim)) + [rng.uniform(0.5, 5), a1])
        U = oracle.prim2cons(P, W)
        W2 = oracle.cons2prim(P, U)
        assert np.allclose(W2, W, rtol=1e-13, atol=1e-14)


def test_nasg_with_zero_covolume_equals_stiffened_gas():
    """NASG with b = q = 0 is the stiffened gas, bit for bit (include/EquationOfState_NASG.H:22-24)."""
    Psg = make_params(2, gamma=(4.4, 1.4), pinf=(6e8, 0.0), eos=1)
    Pn = make_params(2, gamma=(4.4, 1.4), pinf=(6e8, 0.0), eos=3, b=(0.0, 0.0), q=(0.0, 0.0))
    U = np.array([500.0, 0.6, 10.0, -3.0, 3.0e9, 0.5])
    assert np.array_equal(oracle.cons2prim(Psg, U), oracle.cons2prim(Pn, U))


def test_nasg_covolume_enters_pressure():
    """NASG: rho e = (1 - b rho)(p + gamma pinf)/(gamma - 1) + rho q for a pure phase."""
    b, q, g, pinf = 6.7e-4, -1.167e6, 1.19, 6.2178e8
    P = make_params(2, gamma=(g, 1.4), pinf=(pinf, 0.0), eos=3, b=(b, 0.0), q=(q, 0.0))
    rho, p = 900.0, 1.0e7
    rhoe = (1 - b * rho) * (p + g * pinf) / (g - 1) + rho * q
    U = np.array([rho, 0.0, 0.0, 0.0, rhoe, 1.0])
    assert oracle.cons2prim(P, U)[4] == pytest.approx(p, rel=1e-9)


def test_pure_phase_sound_speed():
    """c^2 = gamma (p + pinf)/rho for one stiffened-gas phase, both K variants."""
    P = sg_params()
    rho, p = 1000.0, 1.0e5
    for k in (0, 1):
        c = oracle.sound_speed(P, p, 1.0, rho, 0.0, k)
        assert c == pytest.approx(math.sqrt(4.4 * (p + 6.0e8) / rho), rel=1e-14)


def _pressure_of(P, rho, rhoe, Y1, a1):
    U = np.array([Y1 * rho, (1 - Y1) * rho, 0.0, 0.0, rhoe, a1])
    return oracle.cons2prim(P, U)[4]


@pytest.mark.parametrize("a1", [0.1, 0.5, 0.93])
def test_frozen_mixture_sound_speed_is_thermodynamic(a1):
    """K = 0 (alpha and mass fractions frozen): c^2 = (dp/drho)_e + (h)(dp/d(rho e))_rho with
    h = (rho e + p)/rho, by central finite differences of the discrete EOS."""
    P = make_params(2, gamma=(4.4, 1.4), pinf=(6.0e8, 0.0))
    rho1, rho2, p = 1000.0, 1.2, 1.0e6
    rho = a1 * rho1 + (1 - a1) * rho2
    Y1 = a1 * rho1 / rho
    U, _ = state(2, a1 * rho1, (1 - a1) * rho2, (0.0, 0.0), p, a1, P)
    rhoe = U[4]
    hr, he = rho * 1e-6, rhoe * 1e-6
    dp_drho = (_pressure_of(P, rho + hr, rhoe, Y1, a1) - _pressure_of(P, rho - hr, rhoe, Y1, a1)) / (2 * hr)
    dp_de = (_pressure_of(P, rho, rhoe + he, Y1, a1) - _pressure_of(P, rho, rhoe - he, Y1, a1)) / (2 * he)
    c2 = dp_drho + (rhoe + p) / rho * dp_de
    c = oracle.sound_speed(P, p, a1, a1 * rho1, (1 - a1) * rho2, 0)
    assert c * c == pytest.approx(c2, rel=1e-6)


@pytest.mark.parametrize("a1", [0.2, 0.5, 0.8])
def test_wood_sound_speed_and_kapila_coefficient(a1):
    """K = 1: 1/(rho c^2) = sum alpha_k/(rho_k c_k^2) (Wood) and alpha1 K1 equals the Kapila term
    alpha1 alpha2 (rho2c2^2 - rho1c1^2)/(alpha1 rho2c2^2 + alpha2 rho1c1^2)."""
    P = make_params(2, gamma=(4.4, 1.4), pinf=(6.0e8, 0.0), source_term=1)
    rho1, rho2, p = 1000.0, 1.2, 1.0e6
    r1c1, r2c2 = 4.4 * (p + 6.0e8), 1.4 * p
    rho = a1 * rho1 + (1 - a1) * rho2
    wood = 1.0 / (a1 / r1c1 + (1 - a1) / r2c2)
    c = oracle.sound_speed(P, p, a1, a1 * rho1, (1 - a1) * rho2, 1)
    assert c * c * rho == pytest.approx(wood, rel=1e-13)
    K = oracle.nonconservative_K(P, p, a1, a1 * rho1, (1 - a1) * rho2, 1)
    kap = a1 * (1 - a1) * (r2c2 - r1c1) / (a1 * r2c2 + (1 - a1) * r1c1)
    assert a1 * K[0] == pytest.approx(kap, rel=1e-12)
    assert np.all(oracle.nonconservative_K(P, p, a1, 1.0, 1.0, 0) == 0.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_physical_flux_is_euler_flux(dim):
    """Single phase (alpha1 = 1): [rho u, rho u u + p e_d, (E + p) u, alpha u]."""
    P = make_params(dim, gamma=(1.4, 1.6))
    rho, p = 1.3, 2.1
    vel = [0.7, -0.4, 0.9][:dim]
    U, _ = state(dim, rho, 0.0, vel, p, 1.0, P)
    E = p / 0.4 + 0.5 * rho * sum(v * v for v in vel)
    for d in range(dim):
        F = oracle.cons2flux(P, U, d)
        un = vel[d]
        assert F[0] == pytest.approx(rho * un, rel=1e-15)
        assert F[1] == 0.0
        for m in range(dim):
            assert F[2 + m] == pytest.approx(rho * vel[m] * un + (p if m == d else 0.0), rel=1e-14)
        assert F[2 + dim] == pytest.approx((E + p) * un, rel=1e-14)
        assert F[3 + dim] == pytest.approx(un, rel=1e-15)


# ---- scalar reconstruction -----------------------------------------------------------------------

def _cell_avg_poly(coef, i, h=1.0):
    """Cell average of sum_k coef_k x^k over [i - 1/2, i + 1/2] * h."""
    a, b = (i - 0.5) * h, (i + 0.5) * h
    s = 0.0
    for k, ck in enumerate(coef):
        s += ck * (b ** (k + 1) - a ** (k + 1)) / (k + 1)
    return s / h


@pytest.mark.parametrize("coef", [(1.0,), (2.0, -3.0), (0.5, 1.5, -2.0), (-1.0, 0.3, 0.7)])
def test_weno5_exact_for_quadratics(coef):
    """Each 3-cell candidate reproduces quadratics from cell averages, so any convex weights do."""
    v = [_cell_avg_poly(coef, i) for i in range(-2, 3)]
    q = lambda x: sum(ck * x ** k for k, ck in enumerate(coef))
    assert oracle.weno5(v, 1) == pytest.approx(q(0.5), abs=1e-13)
    assert oracle.weno5(v, -1) == pytest.approx(q(-0.5), abs=1e-13)


def test_weno5_fifth_order_on_smooth_data():
    """Face values of sin from cell averages converge at about 5th order away from extrema."""
    errs = []
    for h in (0.02, 0.01, 0.005):
        x0 = 0.3
        v = [(math.cos((x0 + (i - 0.5) * h)) - math.cos(x0 + (i + 0.5) * h)) / h for i in range(-2, 3)]
        errs.append(abs(oracle.weno5(v, 1) - math.sin(x0 + 0.5 * h)))
    o1, o2 = math.log2(errs[0] / errs[1]), math.log2(errs[1] / errs[2])
    assert o1 > 4.5 and o2 > 4.5, (o1, o2)


def test_weno5_eno_at_a_jump_and_nan_branch():
    """A jump inside the downwind candidate gets zero weight (IS ratio inf); the two smooth
    candidates have IS = 0 and their 0/0 ratio is taken as 1: the face value is exactly 0."""
    assert oracle.weno5([0.0, 0.0, 0.0, 0.0, 1.0], 1) == 0.0
    f, w = oracle.weno5_weights([0.0, 0.0, 0.0, 0.0, 1.0], 1)
    assert w[0] == 0.0 and w[1] == pytest.approx(6.0 / 7.0) and w[2] == pytest.approx(1.0 / 7.0)
    # all smoothness indicators zero: linear weights
    f, w = oracle.weno5_weights([2.0] * 5, 1)
    assert np.allclose(w, [0.3, 0.6, 0.1], rtol=1e-15)
    f, w = oracle.weno5_weights([2.0] * 5, -1)
    assert np.allclose(w, [0.1, 0.6, 0.3], rtol=1e-15)


def test_weno5_mirror_symmetry():
    rng = np.random.default_rng(2)
    for _ in range(100):
        v = rng.normal(size=5)
        assert oracle.weno5(v, 1) == pytest.approx(oracle.weno5(v[::-1], -1), rel=1e-12, abs=1e-14)


def _textbook_slope(a, b, c, lim):
    if lim == 1:   # minmod
        return 0.0 if a * b <= 0 else (a if abs(a) < abs(b) else b)
    if lim == 2:   # van Leer (harmonic mean)
        return 2 * a * b / (a + b) if a * b > 0 else 0.0
    if lim == 3:   # superbee (Roe): max(min(2|a|, |b|), min(|a|, 2|b|)) with the common sign
        if a * b <= 0:
            return 0.0
        s = 1.0 if a > 0 else -1.0
        return s * max(min(2 * abs(a), abs(b)), min(abs(a), 2 * abs(b)))
    if lim == 4:   # monotonized central
        if a * b <= 0:
            return 0.0
        s = 1.0 if a > 0 else -1.0
        return s * min(abs(c), 2 * abs(a), 2 * abs(b))


@pytest.mark.parametrize("lim", [1, 2, 3, 4])
def test_muscl_limiters_are_textbook(lim):
    rng = np.random.default_rng(lim)
    dx = 0.1
    for _ in range(200):
        fm, fc, fp = rng.normal(size=3)
        a, b, c = (fc - fm) / dx, (fp - fc) / dx, (fp - fm) / (2 * dx)
        s = _textbook_slope(a, b, c, lim)
        assert oracle.muscl(fm, fc, fp, dx, 1, lim) == pytest.approx(fc + 0.5 * s * dx, rel=1e-12, abs=1e-14)
        assert oracle.muscl(fm, fc, fp, dx, -1, lim) == pytest.approx(fc - 0.5 * s * dx, rel=1e-12, abs=1e-14)


# ---- Riemann solvers -----------------------------------------------------------------------------

@pytest.mark.parametrize("riemann", ["HLL", "HLLC"])
@pytest.mark.parametrize("dim", [2, 3])
def test_riemann_consistency(riemann, dim):
    """F(U, U) = F(U) and u*_n = u_n."""
    P = make_params(dim, gamma=(4.4, 1.4), pinf=(6e8, 0.0), riemann=riemann)
    U, W = state(dim, 400.0, 0.7, [12.0, -5.0, 3.0][:dim], 2e6, 0.4, P)
    for d in range(dim):
        F, vel, cmax = oracle.riemann(P, U, U, d)
        assert np.allclose(F, oracle.cons2flux(P, U, d), rtol=1e-13, atol=1e-9)
        assert vel[d] == pytest.approx(W[2 + d], rel=1e-13)
        c = oracle.sound_speed(P, W[2 + dim], W[3 + dim], W[0], W[1], 0)
        assert cmax == pytest.approx(c + math.sqrt(sum(x * x for x in W[2:2 + dim])), rel=1e-15)


def test_hllc_resolves_stationary_contact():
    """A material interface at rest (equal p, u = 0) is a steady HLLC solution: no mass or alpha
    flux, momentum flux = p (HLL smears it instead)."""
    P = make_params(2, gamma=(1.4, 1.6), riemann="HLLC")
    UL, _ = state(2, 0.9, 0.5, (0.0, 0.2), 1.0, 0.9, P)
    UR, _ = state(2, 0.1, 9.0, (0.0, -0.3), 1.0, 0.1, P)
    F, vel, _ = oracle.riemann(P, UL, UR, 0)
    assert abs(F[0]) < 1e-15 and abs(F[1]) < 1e-14 and abs(F[5]) < 1e-15
    assert F[2] == pytest.approx(1.0, rel=1e-14) and abs(F[4]) < 1e-14
    assert abs(vel[0]) < 1e-15
    Ph = make_params(2, gamma=(1.4, 1.6), riemann="HLL")
    assert abs(oracle.riemann(Ph, UL, UR, 0)[0][0]) > 1e-3


def _toro_hllc(rhoL, uL, vL, pL, rhoR, uR, vR, pR, gamma):
    """Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, 3rd ed., eqs. 10.26,
    10.37-10.39, with the wave speeds of the reference (arithmetic mean / Davis-type)."""
    cL, cR = math.sqrt(gamma * pL / rhoL), math.sqrt(gamma * pR / rhoR)
    EL = pL / (gamma - 1) + 0.5 * rhoL * (uL * uL + vL * vL)
    ER = pR / (gamma - 1) + 0.5 * rhoR * (uR * uR + vR * vR)
    SL = min(0.5 * (uL + uR) - 0.5 * (cL + cR), uL - cL)
    SR = max(0.5 * (uL + uR) + 0.5 * (cL + cR), uR + cR)
    Ss = (pR - pL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR)) / (rhoL * (SL - uL) - rhoR * (SR - uR))

    def F(rho, u, v, p, E):
        return np.array([rho * u, rho * u * u + p, rho * u * v, (E + p) * u])

    def Us(rho, u, v, p, E, S):
        f = rho * (S - u) / (S - Ss)
        return f * np.array([1.0, Ss, v, E / rho + (Ss - u) * (Ss + p / (rho * (S - u)))])

    UL = np.array([rhoL, rhoL * uL, rhoL * vL, EL])
    UR = np.array([rhoR, rhoR * uR, rhoR * vR, ER])
    if SL >= 0:
        return F(rhoL, uL, vL, pL, EL)
    if Ss >= 0:
        return F(rhoL, uL, vL, pL, EL) + SL * (Us(rhoL, uL, vL, pL, EL, SL) - UL)
    if SR > 0:
        return F(rhoR, uR, vR, pR, ER) + SR * (Us(rhoR, uR, vR, pR, ER, SR) - UR)
    return F(rhoR, uR, vR, pR, ER)


@pytest.mark.parametrize("case", [
    (1.0, 0.0, 0.0, 1.0, 0.125, 0.0, 0.0, 0.1),     # Sod
    (1.0, -2.0, 0.3, 0.4, 1.0, 2.0, -0.1, 0.4),     # 123 problem
    (1.0, 0.0, 0.0, 1000.0, 1.0, 0.0, 0.0, 0.01),   # strong blast
    (5.99924, 19.5975, 0.0, 460.894, 5.99242, -6.19633, 0.0, 46.095),  # shock collision
    (1.0, 3.0, 1.0, 1.0, 0.5, 2.5, 0.0, 0.8),       # supersonic to the right
    (0.8, -2.5, 0.2, 0.7, 1.0, -3.0, 0.0, 1.0),     # supersonic to the left
])
def test_hllc_matches_toro_single_phase(case):
    """Single-phase limit (alpha1 = 1, no phase-2 mass) against Toro's HLLC."""
    rhoL, uL, vL, pL, rhoR, uR, vR, pR = case
    P = make_params(2, gamma=(1.4, 1.67), riemann="HLLC")
    UL, _ = state(2, rhoL, 0.0, (uL, vL), pL, 1.0, P)
    UR, _ = state(2, rhoR, 0.0, (uR, vR), pR, 1.0, P)
    F, vel, _ = oracle.riemann(P, UL, UR, 0)
    T = _toro_hllc(rhoL, uL, vL, pL, rhoR, uR, vR, pR, 1.4)
    got = np.array([F[0], F[2], F[3], F[4]])
    assert np.allclose(got, T, rtol=1e-12, atol=1e-12 * np.max(np.abs(T)))
    assert F[1] == 0.0
    # the alpha flux is alpha * u*: the face velocity of the non-conservative term is consistent
    assert F[5] == pytest.approx(vel[0], rel=1e-13, abs=1e-15)


def test_hll_matches_textbook():
    """HLL (Harten-Lax-van Leer) with S- = min(0, uL-cL, uR-cR), S+ = max(0, uL+cL, uR+cR)."""
    P = make_params(2, gamma=(1.4, 1.67), riemann="HLL")
    rhoL, uL, pL, rhoR, uR, pR = 1.0, 0.3, 1.0, 0.125, -0.2, 0.1
    UL, _ = state(2, rhoL, 0.0, (uL, 0.0), pL, 1.0, P)
    UR, _ = state(2, rhoR, 0.0, (uR, 0.0), pR, 1.0, P)
    F, _, _ = oracle.riemann(P, UL, UR, 0)
    cL, cR = math.sqrt(1.4 * pL / rhoL), math.sqrt(1.4 * pR / rhoR)
    Sm, Sp = min(0.0, uL - cL, uR - cR), max(0.0, uL + cL, uR + cR)
    FL, FR = oracle.cons2flux(P, UL, 0), oracle.cons2flux(P, UR, 0)
    ref = (Sp * FL - Sm * FR + Sp * Sm * (UR - UL)) / (Sp - Sm)
    assert np.allclose(F, ref, rtol=1e-13, atol=1e-14)
    # upwind when supersonic
    UL2, _ = state(2, 1.0, 0.0, (5.0, 0.0), 1.0, 1.0, P)
    UR2, _ = state(2, 0.5, 0.0, (4.0, 0.0), 0.5, 1.0, P)
    assert np.allclose(oracle.riemann(P, UL2, UR2, 0)[0], oracle.cons2flux(P, UL2, 0), rtol=1e-14)
