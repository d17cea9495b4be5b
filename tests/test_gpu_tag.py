"""GPU refinement tagging (cmp_level_tag through the C ABI) against the oracle on the same
ghost-filled states: the tags of every valid cell are identical (2D and 3D, two-phase and N-phase
models, gradient / value criteria on state components and on the derived rho, P, u)."""
import numpy as np
import pytest

import oracle
from compas_amd import Layout, Level, cases

pytestmark = pytest.mark.gpu

CRITERIA = [[(None, "grad", 0.02)], [("P", "grad", 0.003)], [(None, "value_range", 0.01, 0.99)],
            [("rho", "value_greater", 1.5), ("u", "value_less", 0.2), (None, "grad", 0.05)]]


@pytest.mark.parametrize("dim,nphase", [(2, 0), (3, 0), (2, 3)])
@pytest.mark.parametrize("crit", range(len(CRITERIA)))
def test_tags_match_oracle(gpu, dim, nphase, crit):
    import torch
    n = 32 if dim == 2 else 16
    if nphase:
        c = cases.nphase_smooth(dim, n, seed=4, nphase=nphase, periodic=False, max_grid_size=n // 2)
        alpha = 2 * nphase + dim
    else:
        c = cases.random_smooth(dim, n, seed=4, periodic=False, max_grid_size=n // 2)
        alpha = 3 + dim
    P, G = c["params"], c["geom"]
    criteria = [(alpha if v is None else v,) + tuple(rest) for v, *rest in CRITERIA[crit]]
    lay = Layout(dim, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    oracle.level_fill_ghosts(P, G, lay, buf)
    ref = oracle.level_tag(P, lay, buf, criteria)
    U = torch.from_numpy(buf).to(gpu)
    tags = torch.full((ref.size,), 7, dtype=torch.uint8, device=gpu)
    lev = Level(P, G, lay)
    lev.ErrorEst(U, criteria, tags)
    assert lev.check_errors() == 0
    lev.close()
    got = tags.cpu().numpy()
    assert np.array_equal(got, ref), (int((got != ref).sum()), ref.sum())
    assert ref.sum() > 0


def test_regrid_from_gpu_tags(gpu):
    """Regrid of the three-phase rings at 128^2: GPU tags on every volume-fraction gradient and on
    |grad P| (the AMR configuration's criteria), clustered into blocking-factor-aligned boxes of at
    most 32 cells that cover every tag, refined by 2 for the next level."""
    import torch
    from compas_amd import regrid
    c = cases.nphase_rings_2d(128, boxes=None)
    P, G = c["params"], c["geom"]
    lay = Layout(2, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    oracle.level_fill_ghosts(P, G, lay, buf)
    crit = [(6, "grad", 0.1), (7, "grad", 0.1), (8, "grad", 0.1), ("P", "grad", 0.1)]
    lev = Level(P, G, lay)
    U = torch.from_numpy(buf).to(gpu)
    boxes, tags = regrid.regrid(lev, U, crit, (128, 128), ratio=2, blocking_factor=8, max_grid_size=32)
    lev.close()
    assert np.array_equal(tags.astype(bool), regrid.tags_to_global(lay, oracle.level_tag(P, lay, buf, crit),
                                                                   (128, 128)).astype(bool))
    assert tags.sum() > 0 and len(boxes) > 0
    cover = np.zeros((256, 256), dtype=int)
    for lo, hi in boxes:
        assert all(l % 16 == 0 for l in lo) and all(h - l + 1 <= 64 for l, h in zip(lo, hi))
        cover[lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert cover.max() == 1
    fine_tags = np.kron(tags, np.ones((2, 2), dtype=np.uint8)).astype(bool)
    assert np.all(cover[fine_tags] == 1)
