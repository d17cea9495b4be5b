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
