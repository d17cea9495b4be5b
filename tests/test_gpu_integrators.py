"""GPU ForwardEuler, TVD-RK2 and RK4 steps (cmp_level_rk_step through the C ABI) against the oracle's
on a smooth 2D two-phase case and a 3D case: per-variable max difference relative to the variable's
largest magnitude at most 1e-12 after one step; integrator 3 is the fused TVD-RK3 step."""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, Level, cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,integrator", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 4)])
def test_rk_step_matches_oracle(gpu, dim, integrator):
    import torch
    c = cases.random_smooth(dim, 32 if dim == 2 else 16, 7, max_grid_size=16 if dim == 2 else 8)
    P, G = c["params"], c["geom"]
    lay = Layout(dim, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    dt = 1.5e-3
    ref = buf.copy()
    oracle.level_rk_step(P, G, lay, ref, dt, integrator)
    Un = torch.from_numpy(buf).to(gpu)
    U1, U2, U3 = (torch.zeros_like(Un) for _ in range(3))
    dtd = torch.full((1,), dt, dtype=torch.float64, device=gpu)
    cm = torch.zeros(1, dtype=torch.float64, device=gpu)
    lev = Level(P, G, lay)
    lev.rk_step(integrator, Un, U1, U2, dtd, cm, U3=U3)
    assert lev.check_errors() == 0
    lev.close()
    got, want = lay.unpack(Un.cpu().numpy(), c["shape"]), lay.unpack(ref, c["shape"])
    err = np.abs(got - want).reshape(P.nstate, -1).max(axis=1) / np.abs(want).reshape(P.nstate, -1).max(axis=1)
    assert err.max() <= 1e-12, err
    assert cm.item() > 0.0


@pytest.mark.parametrize("integrator,with_pf", [(1, False), (2, False), (4, False), (2, True), (4, True)])
def test_advance_with_the_level_integrator_matches_oracle(gpu, integrator, with_pf):
    """cmp_level_advance (graph-replayed steps) and, with a Phase-Field step, its sub-steps under
    cmp_level_set_integrator against the oracle's step loop with the same run.TimeIntegrator."""
    import torch
    from compas_amd import make_pf_params
    c = cases.random_smooth(2, 32, 9, max_grid_size=16)
    P, G = c["params"], c["geom"]
    lay = Layout(2, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    pf = make_pf_params(mechanism="ACDI", explicit_rc=1) if with_pf else None
    nsteps = 6
    ref = buf.copy()
    hist = oracle.level_advance_int(P, G, lay, ref, nsteps, 0.4, integrator, pf=pf)
    Un = torch.from_numpy(buf).to(gpu)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dt = torch.zeros(1, dtype=torch.float64, device=gpu)
    cm = torch.zeros(1, dtype=torch.float64, device=gpu)
    dth = torch.zeros(nsteps, dtype=torch.float64, device=gpu)
    lev = Level(P, G, lay)
    if pf is not None:
        lev.pf_setup(pf)
    lev.set_integrator(integrator)
    lev.advance(Un, U1, U2, nsteps, 0.4, dt, cm, dt_hist=dth, use_graph=True)
    assert lev.check_errors() == 0
    lev.close()
    assert np.all(np.abs(dth.cpu().numpy() - hist) <= 1e-12 * hist)
    got, want = lay.unpack(Un.cpu().numpy(), c["shape"]), lay.unpack(ref, c["shape"])
    err = np.abs(got - want).reshape(P.nstate, -1).max(axis=1) / np.abs(want).reshape(P.nstate, -1).max(axis=1)
    assert err.max() <= 1e-11, err


def test_level_clock_and_stop_time(gpu):
    """The level's time advances by every dt; with stop_time the dt that would pass it (beyond
    stop_time - 1e-3 dt) is cut to end on it, and the next step's limiter sees the cut dt."""
    import torch
    c = cases.random_smooth(2, 32, 9, max_grid_size=16)
    P, G = c["params"], c["geom"]
    lay = Layout(2, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    ref = buf.copy()
    hist = oracle.level_advance(P, G, lay, ref, 4, 0.4)

    def run(stop):
        Un = torch.from_numpy(buf).to(gpu)
        U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
        dt = torch.zeros(1, dtype=torch.float64, device=gpu)
        cm = torch.zeros(1, dtype=torch.float64, device=gpu)
        dth = torch.zeros(4, dtype=torch.float64, device=gpu)
        lev = Level(P, G, lay)
        lev.set_time(0.0, stop)
        lev.advance(Un, U1, U2, 4, 0.4, dt, cm, dt_hist=dth, use_graph=True)
        t = lev.get_time()
        assert lev.check_errors() == 0
        lev.close()
        return dth.cpu().numpy(), t

    h, t = run(-1.0)
    assert np.all(np.abs(h - hist) <= 1e-12 * hist)
    assert t == ((h[0] + h[1]) + h[2]) + h[3]
    stop = h[0] + h[1] + 0.5 * h[2]
    h2, t2 = run(stop)
    assert np.array_equal(h2[:2], h[:2])
    assert h2[2] == stop - (h2[0] + h2[1])
    assert t2 == ((h2[0] + h2[1]) + h2[2]) + h2[3]
    assert abs(((h2[0] + h2[1]) + h2[2]) - stop) <= 1e-15 * stop
    # past stop_time every further dt is cut to stop_time - t (zero-length steps)
    assert h2[3] == stop - ((h2[0] + h2[1]) + h2[2])
