"""GPU tests of the batch of levels (the served-stream form, cmp_batch_* through the C ABI):

- every case advanced in a batch (different grids, box splits, EOS parameters, boundary
  conditions) is bit-identical to the same case advanced alone by the single-level path, with the
  same simulated time; cases admitted and retired mid-run included (a result is the same alone and
  in the mix);
- admission is refused when the arenas have no room (CMP_ERR_NOMEM) and succeeds once a retired
  case has freed its range;
- a batched case against the oracle: two steps of a shock-bubble <= 1e-12 per variable (the
  momenta against the larger momentum component's max: the transverse momentum starts at zero and
  is still small, so its own max is no scale for round-off).
"""
import numpy as np
import pytest

import oracle
from compas_amd import Batch, Layout, Level, cases, make_geometry
from tests._util import maxrel_per_var
from tests.test_gpu_parity import _staggered_boxes

pytestmark = pytest.mark.gpu


def _alone(c, nsteps, pf=None):
    import torch
    dev = torch.device("cuda:0")
    P = c["params"]
    lay = Layout(P.dim, c["boxes"], 3, P.nstate)
    Un = torch.from_numpy(lay.pack(c["U"])).to(dev)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dt = torch.zeros(1, dtype=torch.float64, device=dev)
    cm = torch.zeros(1, dtype=torch.float64, device=dev)
    hist = torch.zeros(max(1, nsteps), dtype=torch.float64, device=dev)
    lev = Level(P, c["geom"], lay)
    if pf is not None:
        lev.pf_setup(pf)
    lev.advance(Un, U1, U2, nsteps, c["cfl"], dt, cm, hist, use_graph=False)
    assert lev.check_errors() == 0
    lev.close()
    return lay.unpack(Un.cpu().numpy(), c["shape"]), float(hist.cpu().numpy()[:nsteps].sum())


def _walled(seed):
    n = 40
    c = cases.random_smooth(2, n, seed=seed, periodic=False)
    c["geom"] = make_geometry((n, n), (0.0, 0.0), (1.0, 1.0), (0, 0), lo_bc=[4, 3], hi_bc=[5, 2])
    c["boxes"] = _staggered_boxes(2, n)
    return c


def _mix():
    return [cases.shock_bubble_2d(64, 64, 2.2, 0.2, 1.6, 1.4, 0.2, max_grid_size=32),
            cases.shock_bubble_2d(128, 32, 1.5, 5.0, 1.2, 1.67, 0.15, max_grid_size=32),
            cases.random_smooth(2, 48, seed=3, max_grid_size=16),
            _walled(7)]


def _admit(b, c):
    P = c["params"]
    lay = Layout(P.dim, c["boxes"], 3, P.nstate)
    return b.admit(P, c["geom"], lay, lay.pack(c["U"]), c["cfl"]), lay


def test_batched_cases_equal_alone_runs(gpu):
    cs = _mix()
    b = Batch(cs[0]["params"], 4_000_000, 8)
    ids = []
    for c in cs:
        cid, lay = _admit(b, c)
        assert cid is not None
        ids.append((cid, lay))
    for _ in range(6):
        b.step()
    for c, (cid, lay) in zip(cs, ids):
        assert b.check_errors(cid) == 0
        buf, t, k = b.state(cid)
        ref, tref = _alone(c, 6)
        assert k == 6
        assert t == tref
        assert np.array_equal(lay.unpack(buf, c["shape"]), ref)
    n, ntiles, free = b.info()
    assert n == 4 and ntiles > 0
    b.close()


def test_admission_and_retirement_mid_run(gpu):
    cs = _mix()
    b = Batch(cs[0]["params"], 4_000_000, 8)
    a, la = _admit(b, cs[0])
    bb, lb = _admit(b, cs[1])
    b.step()
    b.step()
    c, lc = _admit(b, cs[3])
    b.step()
    b.step()
    b.retire(a)
    d, ld = _admit(b, cs[2])
    for _ in range(2):
        b.step()
    for cid, case, lay, nsteps in ((bb, cs[1], lb, 6), (c, cs[3], lc, 4), (d, cs[2], ld, 2)):
        buf, t, k = b.state(cid)
        ref, tref = _alone(case, nsteps)
        assert k == nsteps and t == tref
        assert np.array_equal(lay.unpack(buf, case["shape"]), ref)
    b.close()


def test_admission_waits_for_memory(gpu):
    c0 = cases.shock_bubble_2d(64, 64, 2.0, 0.5, 1.4, 1.4, 0.2, max_grid_size=32)
    lay = Layout(2, c0["boxes"], 3, 6)
    per_case = lay.total
    b = Batch(c0["params"], 2 * per_case + per_case // 2, 8)
    ids = [_admit(b, c0)[0] for _ in range(3)]
    assert ids[0] is not None and ids[1] is not None and ids[2] is None
    b.step()
    b.retire(ids[0])
    again, _ = _admit(b, c0)
    assert again is not None
    b.step()
    buf, t, k = b.state(again)
    ref, tref = _alone(c0, 1)
    assert k == 1 and np.array_equal(lay.unpack(buf, c0["shape"]), ref)
    buf, t, k = b.state(ids[1])
    assert k == 2
    b.close()


def test_batched_case_matches_oracle(gpu):
    c = cases.shock_bubble_2d(96, 48, 2.0, 0.3, 1.5, 1.4, 0.15, max_grid_size=32)
    P = c["params"]
    b = Batch(P, 1_000_000, 2)
    cid, lay = _admit(b, c)
    b.step()
    b.step()
    buf, t, k = b.state(cid)
    b.close()
    ref_buf = lay.pack(c["U"])
    hist = oracle.level_advance(P, c["geom"], lay, ref_buf, 2, c["cfl"])
    assert abs(t - hist.sum()) <= 1e-14 * t
    got, ref = lay.unpack(buf, c["shape"]), lay.unpack(ref_buf, c["shape"])
    rel = maxrel_per_var(got, ref)
    mom = max(np.max(np.abs(ref[2])), np.max(np.abs(ref[3])))
    rel[2:4] = [np.max(np.abs(got[q] - ref[q])) / mom for q in (2, 3)]
    assert rel.max() <= 1e-12, rel


def test_served_stream_results_equal_alone_runs(gpu):
    """The request loop end to end: eight seeded cases of three grid sizes arriving at 2000 cases/s
    (arrivals spread over a few milliseconds, so that several are in flight together and later ones
    join mid-run; how many depends on the GPU's speed), 12 steps each; every final state is
    bit-identical to the case run alone."""
    from compas_amd import serve
    grids = [(64, 64), (128, 32), (96, 48)]
    res = serve.serve(2000.0, 8, steps=12, seed=5, max_grid_size=32, grids=grids, keep=set(range(8)))
    assert res["mean_cases_per_tick"] > 1.0
    assert res["p99_s"] >= res["p50_s"] > 0.0
    for i in range(8):
        c = cases.served_case(5 * 100003 + i, max_grid_size=32, grids=grids)
        lay = Layout(2, c["boxes"], 3, 6)
        ref, _ = _alone(c, 12)
        assert np.array_equal(lay.unpack(res["finals"][i], c["shape"]), ref), i


@pytest.mark.parametrize("mech,rc", [("CAC-Adv", 0), ("ACDI", 1)])
def test_served_stream_with_phase_field_equals_alone_runs(gpu, mech, rc):
    """The request loop with the Phase-Field step of every case after its flow step (the cases'
    steps concurrent on their own streams): six seeded cases of three grid sizes arriving at 2000
    cases/s, 6 steps each; every final state is bit-identical to the case run alone through the
    single-level path with the same Phase-Field step, and the interface has moved."""
    from compas_amd import make_pf_params, serve
    pf = make_pf_params(mechanism=mech, explicit_rc=rc, cfl=0.3, eta_multiplier=0.5)
    grids = [(64, 64), (128, 32), (96, 48)]
    res = serve.serve(2000.0, 6, steps=6, seed=9, max_grid_size=32, grids=grids, keep=set(range(6)), pf=pf)
    assert res["p99_s"] >= res["p50_s"] > 0.0
    for i in range(6):
        c = cases.served_case(9 * 100003 + i, max_grid_size=32, grids=grids)
        lay = Layout(2, c["boxes"], 3, 6)
        ref, _ = _alone(c, 6, pf)
        flow, _ = _alone(c, 6)
        got = lay.unpack(res["finals"][i], c["shape"])
        assert np.array_equal(got, ref), (i, np.max(np.abs(got - ref)))
        assert np.max(np.abs(ref[5] - flow[5])) > 1e-9, i
