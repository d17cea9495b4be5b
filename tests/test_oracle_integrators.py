"""The oracle's other per-level integrators (run.TimeIntegrator = ForwardEuler, TVD-RK2, RK4), pinned
by the textbook schemes assembled from the oracle's right-hand side (a filled copy per stage) and by
their temporal order of accuracy on a smooth periodic case."""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, cases


def _case():
    c = cases.random_smooth(2, 16, 11, max_grid_size=8)
    P, G = c["params"], c["geom"]
    lay = Layout(2, c["boxes"], 3, P.nstate)
    return P, G, lay, lay.pack(c["U"]), c["shape"]


def _L(P, G, lay, U):
    b = U.copy()
    oracle.level_fill_ghosts(P, G, lay, b)
    return oracle.level_rhs(P, G, lay, b)[0]


def _valid(lay, buf, shape):
    return lay.unpack(buf, shape)


def test_forward_euler_and_tvd_rk2_and_rk4_are_the_textbook_schemes():
    P, G, lay, U0, shape = _case()
    dt = 2.0e-3
    k1 = _L(P, G, lay, U0)
    fe = U0 + dt * k1
    got = U0.copy()
    oracle.level_rk_step(P, G, lay, got, dt, 1)
    assert np.allclose(_valid(lay, got, shape), _valid(lay, fe, shape), rtol=1e-14, atol=1e-14)
    # TVD-RK2 = mean of U^n and two forward-Euler steps
    u1 = U0 + dt * k1
    rk2 = 0.5 * U0 + 0.5 * (u1 + dt * _L(P, G, lay, u1))
    got = U0.copy()
    oracle.level_rk_step(P, G, lay, got, dt, 2)
    assert np.allclose(_valid(lay, got, shape), _valid(lay, rk2, shape), rtol=1e-13, atol=1e-13)
    # classical RK4
    k2 = _L(P, G, lay, U0 + 0.5 * dt * k1)
    k3 = _L(P, G, lay, U0 + 0.5 * dt * k2)
    k4 = _L(P, G, lay, U0 + dt * k3)
    rk4 = U0 + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    got = U0.copy()
    oracle.level_rk_step(P, G, lay, got, dt, 4)
    assert np.allclose(_valid(lay, got, shape), _valid(lay, rk4, shape), rtol=1e-13, atol=1e-13)
    # 3 is the TVD-RK3 step
    a, b = U0.copy(), U0.copy()
    oracle.level_rk_step(P, G, lay, a, dt, 3)
    oracle.level_rk3_step(P, G, lay, b, dt)
    assert np.array_equal(_valid(lay, a, shape), _valid(lay, b, shape))


@pytest.mark.parametrize("integrator,order", [(1, 1), (2, 2), (3, 3), (4, 4)])
def test_temporal_order(integrator, order):
    P, G, lay, U0, shape = _case()
    T = 8.0e-3
    ref = U0.copy()
    for _ in range(64):
        oracle.level_rk_step(P, G, lay, ref, T / 64, 4)
    errs = []
    for n in (4, 8):
        u = U0.copy()
        for _ in range(n):
            oracle.level_rk_step(P, G, lay, u, T / n, integrator)
        errs.append(np.abs(_valid(lay, u, shape) - _valid(lay, ref, shape)).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > order - 0.3, (errs, rate)


def test_pf_sub_step_integrators_compose_the_pf_rhs():
    """ForwardEuler of the Phase-Field sub-step is U + pf_dt L_pf(U) at the fixed Grad_Q, and RK3
    through the integrator switch is the Phase-Field RK3 sub-step."""
    from compas_amd import FaceLayout, make_pf_params
    P, G, lay, U0, shape = _case()
    pf = make_pf_params(flux_phase2flux=1)
    flay = FaceLayout(2, lay.boxes)
    gq, gm, _ = oracle.level_pf_gradq_step(P, pf, G, lay, U0.copy(), 0.01, flay.total)
    pf_dt = 0.4 * G.dx[0] / gm
    b = U0.copy()
    oracle.level_fill_ghosts(P, G, lay, b)
    Lpf = oracle.level_pf_rhs(P, pf, G, lay, b, gq)
    fe = U0 + pf_dt * (Lpf[0] if isinstance(Lpf, tuple) else Lpf)
    got = U0.copy()
    oracle.level_pf_rk_step(P, pf, G, lay, got, gq, pf_dt, 1)
    assert np.allclose(_valid(lay, got, shape), _valid(lay, fe, shape), rtol=1e-14, atol=1e-14)
    a, c = U0.copy(), U0.copy()
    oracle.level_pf_rk_step(P, pf, G, lay, a, gq, pf_dt, 3)
    oracle.level_pf_rk3_step(P, pf, G, lay, c, gq, pf_dt)
    assert np.array_equal(a, c)
