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


def _clock_run(gpu, c, lay, nsteps, stop, pf=None, use_graph=True):
    import torch
    P, G = c["params"], c["geom"]
    Un = torch.from_numpy(lay.pack(c["U"])).to(gpu)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dt = torch.zeros(1, dtype=torch.float64, device=gpu)
    cm = torch.zeros(1, dtype=torch.float64, device=gpu)
    dth = torch.zeros(nsteps, dtype=torch.float64, device=gpu)
    lev = Level(P, G, lay)
    if pf is not None:
        lev.pf_setup(pf)
    lev.set_time(0.0, stop)
    lev.advance(Un, U1, U2, nsteps, c["cfl"], dt, cm, dt_hist=dth, use_graph=use_graph)
    t = lev.get_time()
    assert lev.check_errors() == 0
    lev.close()
    return Un.cpu().numpy(), dth.cpu().numpy(), t


@pytest.mark.parametrize("use_graph", [False, True])
def test_stop_time_inside_a_run_with_the_phase_field_step(gpu, use_graph):
    """stop_time half-way through the third step of a five-step run with the Phase-Field step
    (ACDI, explicit RC): the third step is cut to end on stop_time, the two steps after it have
    dt = 0 and change nothing (no flow stage, no Phase-Field step), the run reports no error, and its
    state equals, bit for bit, a run of three steps to the same stop_time."""
    from compas_amd import make_pf_params
    c = cases.advected_interface_2d(n=32, max_grid_size=16)
    lay = Layout(2, c["boxes"], 3, 6)
    pf = make_pf_params(eta_multiplier=0.5, cfl=0.3)
    _, h, _ = _clock_run(gpu, c, lay, 3, -1.0, pf, use_graph)
    stop = (h[0] + h[1]) + 0.5 * h[2]
    U5, h5, t5 = _clock_run(gpu, c, lay, 5, stop, pf, use_graph)
    U3, h3, t3 = _clock_run(gpu, c, lay, 3, stop, pf, use_graph)
    assert np.array_equal(h5[:2], h[:2]) and h5[2] == stop - (h5[0] + h5[1])
    assert h5[3] == 0.0 and h5[4] == 0.0
    assert t5 == stop and t3 == stop
    assert np.array_equal(U5, U3)
    # the cut step moved the state
    U2, _, _ = _clock_run(gpu, c, lay, 2, -1.0, pf, use_graph)
    assert not np.array_equal(U3, U2)


def test_stop_time_cut_can_lengthen_a_step(gpu):
    """With the default run.do_subcycle = 1, ComputeDt's cut (t + dt > stop_time - 1e-3 dt gives
    dt = stop_time - t, source/COMPAS.cpp:2278-2283) also applies when stop_time lies up to 1e-3 dt
    beyond t + dt: the step is lengthened to end on stop_time."""
    c = cases.random_smooth(2, 32, 9, max_grid_size=16)
    c["cfl"] = 0.4
    lay = Layout(2, c["boxes"], 3, c["params"].nstate)
    _, h, _ = _clock_run(gpu, c, lay, 3, -1.0)
    stop = (h[0] + h[1]) + h[2] * (1.0 + 5.0e-4)
    _, h2, t2 = _clock_run(gpu, c, lay, 3, stop)
    assert np.array_equal(h2[:2], h[:2])
    assert h2[2] == stop - (h2[0] + h2[1]) and h2[2] > h[2]
    assert t2 == stop
