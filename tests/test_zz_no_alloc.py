"""The allocation-free level right-hand side, cmp_level_rhs_noalloc, through the C ABI against the
entry point it stands in for (cmp_level_create + cmp_level_rhs) on the same ghost-filled state.

Both run the same stage kernel over the same tile work-list, so the bar is bit-identical dUdt and
c_max. Two calls in a row, the second with a smaller input, check that the fixed device tables of
the first call leave no trace in the second; sizes beyond the fixed tables come back as
CMP_ERR_SCRATCH before any GPU work.
"""
import ctypes

import numpy as np
import pytest

from compas_amd import Layout, Level, capi, cases, scheme_ghosts

pytestmark = pytest.mark.gpu


def _level_rhs_both(gpu, c, scheme):
    import torch
    P, G = c["params"], c["geom"]
    ng = scheme_ghosts(scheme)
    lay = Layout(P.dim, c["boxes"], ng, P.nstate)
    U = torch.from_numpy(lay.pack(c["U"])).to(gpu)
    lev = Level(P, G, lay)
    lev.FillPatch(U)
    ref = torch.zeros_like(U)
    got = torch.zeros_like(U)
    cm_ref = torch.zeros(1, dtype=torch.float64, device=gpu)
    cm_got = torch.zeros(1, dtype=torch.float64, device=gpu)
    err = torch.zeros(1, dtype=torch.int32, device=gpu)
    lev.compute_dUdt_FV(ref, U, cm_ref)
    lib = capi.load()
    print("noalloc rhs: dim %d, %d boxes, %d doubles" % (P.dim, lay.nbox, lay.total), flush=True)
    rc = lib.cmp_level_rhs_noalloc(ctypes.byref(P), ctypes.byref(G), lay.nbox, lay.cboxes, ng,
                                   U.data_ptr(), got.data_ptr(), cm_got.data_ptr(), err.data_ptr(), None)
    assert rc == capi.CMP_OK, capi.ERROR_NAMES.get(rc, rc)
    torch.cuda.synchronize()
    assert lev.check_errors() == 0
    assert int(err.item()) == 0
    assert torch.equal(got, ref)
    assert cm_got.item() == cm_ref.item() and cm_ref.item() > 0.0
    # the valid cells are not all zero: the comparison is not vacuous
    assert float(torch.max(torch.abs(ref))) > 0.0
    lev.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_rhs_equals_level_rhs_twice(gpu, dim):
    big = cases.random_smooth(dim, 96 if dim == 2 else 32, seed=11, max_grid_size=16 if dim == 2 else 8,
                              scheme="WENO5", pinf=(0.0, 0.0))
    _level_rhs_both(gpu, big, "WENO5")
    # second call: fewer, uneven boxes of another size (the first call's tables are longer)
    small = cases.random_smooth(dim, 40 if dim == 2 else 20, seed=12, max_grid_size=24 if dim == 2 else 12,
                                scheme="WENO5", pinf=(0.0, 0.0))
    _level_rhs_both(gpu, small, "WENO5")


def test_noalloc_rhs_rejects_what_does_not_fit(gpu):
    import torch
    lib = capi.load()
    c = cases.random_smooth(2, 16, seed=1, scheme="WENO5", pinf=(0.0, 0.0))
    P, G = c["params"], c["geom"]
    t = torch.zeros(8, dtype=torch.float64, device=gpu)
    # more boxes than the table holds
    n = capi.CMP_NOALLOC_MAX_BOXES + 1
    boxes = (capi.cmp_box * n)()
    rc = lib.cmp_level_rhs_noalloc(ctypes.byref(P), ctypes.byref(G), n, boxes, 3, t.data_ptr(), t.data_ptr(),
                                   None, None, None)
    assert rc == capi.CMP_ERR_SCRATCH
    # more tiles than the work-list holds: an 8192^2 domain in 512^2 boxes. A tile holds at most 256
    # cells whatever its shape, so the level has at least 8192^2 / 256 tiles; the call below must come
    # back before any GPU work (the state buffer is 8 doubles), so the count is checked first.
    from compas_amd import make_geometry, split_domain
    N = 8192
    assert N * N // 256 > capi.CMP_NOALLOC_MAX_TILES
    G2 = make_geometry((N, N), (0, 0), (1, 1), (1, 1))
    lay_boxes = split_domain((N, N), 512)
    assert len(lay_boxes) <= capi.CMP_NOALLOC_MAX_BOXES
    cb = (capi.cmp_box * len(lay_boxes))()
    off = 0
    for b, (lo, hi) in enumerate(lay_boxes):
        for d in range(2):
            cb[b].lo[d], cb[b].hi[d] = lo[d], hi[d]
        cb[b].offset = off
        off += (hi[0] - lo[0] + 7) * (hi[1] - lo[1] + 7) * P.nstate
    rc = lib.cmp_level_rhs_noalloc(ctypes.byref(P), ctypes.byref(G2), len(lay_boxes), cb, 3, t.data_ptr(),
                                   t.data_ptr(), None, None, None)
    assert rc == capi.CMP_ERR_SCRATCH
    assert np.isfinite(t.cpu().numpy()).all()
