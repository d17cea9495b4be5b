"""Regridding on the host: refinement tags -> the next level's box list (this library's own
clustering; the reference hands its TagBoxArray to AMReX's Berger-Rigoutsos clustering,
AmrCore::MakeNewGrids, with amr.n_error_buf, amr.blocking_factor, amr.grid_eff and
amr.max_grid_size).

`cluster(tags, ...)` works on the tags of one level (a global [(k,) j, i] uint8 array):

1. tags grown by n_error_buf cells, then coarsened by the blocking factor (a block is tagged when
   any of its cells is);
2. Berger-Rigoutsos on the blocks: a box whose tagged fraction is below grid_eff is cut, at a
   plane without tags when one exists (the middlemost), else where the second derivative of the
   tag signature changes sign most strongly, else in the middle of its longest side;
3. boxes longer than max_grid_size (in cells) are chopped into near-equal pieces on block lines;
4. the boxes, in cells of the level, cover every tag; `refine` maps them to the next level.

`regrid(level, U, criteria, ...)` runs the GPU tagging (cmp_level_tag) and clusters its tags.
"""
import numpy as np


def _grow(tags, n):
    """tags grown by n cells along each axis in turn (no wrap across the array's ends)."""
    if n <= 0:
        return tags.astype(bool)
    out = tags.astype(bool).copy()
    for ax in range(tags.ndim):
        acc = out.copy()
        for s in range(1, min(n, out.shape[ax] - 1) + 1):
            hi = [slice(None)] * out.ndim
            lo = [slice(None)] * out.ndim
            hi[ax], lo[ax] = slice(s, None), slice(None, -s)
            acc[tuple(hi)] |= out[tuple(lo)]
            acc[tuple(lo)] |= out[tuple(hi)]
        out = acc
    return out


def _coarsen(tags, bf):
    shp = [-(-n // bf) for n in tags.shape]
    pad = [(0, s * bf - n) for s, n in zip(shp, tags.shape)]
    t = np.pad(tags, pad)
    for ax in range(t.ndim):
        t = t.reshape(t.shape[:ax] + (shp[ax], bf) + t.shape[ax + 1:]).any(axis=ax + 1)
    return t


def _split(block, lo, hi, eff, out):
    """Berger-Rigoutsos on the tagged blocks inside [lo, hi] (block indices, numpy axis order)."""
    sl = tuple(slice(l, h + 1) for l, h in zip(lo, hi))
    sub = block[sl]
    if not sub.any():
        return
    # shrink to the tags' bounding box
    nz = [np.nonzero(sub.any(axis=tuple(a for a in range(sub.ndim) if a != ax)))[0] for ax in range(sub.ndim)]
    hi = [l + int(z[-1]) for l, z in zip(lo, nz)]
    lo = [l + int(z[0]) for l, z in zip(lo, nz)]
    sl = tuple(slice(l, h + 1) for l, h in zip(lo, hi))
    sub = block[sl]
    if sub.mean() >= eff or all(h == l for l, h in zip(lo, hi)):
        out.append((lo, hi))
        return
    # a plane without tags, the one nearest the middle
    best = None
    for ax in range(sub.ndim):
        sig = sub.sum(axis=tuple(a for a in range(sub.ndim) if a != ax))
        holes = [c for c in range(1, len(sig) - 1) if sig[c] == 0]
        for c in holes:
            score = abs(c - (len(sig) - 1) / 2.0)
            if best is None or score < best[0]:
                best = (score, ax, c, True)
    if best is None:
        # the strongest sign change of the signature's second difference
        for ax in range(sub.ndim):
            sig = sub.sum(axis=tuple(a for a in range(sub.ndim) if a != ax)).astype(float)
            if len(sig) < 4:
                continue
            lap = sig[:-2] - 2 * sig[1:-1] + sig[2:]
            for c in range(len(lap) - 1):
                if lap[c] * lap[c + 1] < 0:
                    jump = abs(lap[c] - lap[c + 1])
                    if best is None or -jump < best[0]:
                        best = (-jump, ax, c + 2, False)
    if best is None:
        ax = int(np.argmax([h - l for l, h in zip(lo, hi)]))
        best = (0, ax, (hi[ax] - lo[ax] + 1) // 2, False)
    _, ax, c, hole = best
    a_hi = list(hi)
    b_lo = list(lo)
    if hole:
        a_hi[ax] = lo[ax] + c - 1
        b_lo[ax] = lo[ax] + c + 1
    else:
        a_hi[ax] = lo[ax] + c - 1
        b_lo[ax] = lo[ax] + c
    _split(block, lo, a_hi, eff, out)
    _split(block, b_lo, hi, eff, out)


def _chop(lo, hi, maxb):
    """Chop a block box into pieces of at most maxb blocks per side, near-equal sizes."""
    pieces = [(list(lo), list(hi))]
    for ax in range(len(lo)):
        nxt = []
        for l, h in pieces:
            n = h[ax] - l[ax] + 1
            k = -(-n // maxb)
            base, extra = divmod(n, k)
            start = l[ax]
            for q in range(k):
                m = base + (1 if q < extra else 0)
                a, b = list(l), list(h)
                a[ax], b[ax] = start, start + m - 1
                nxt.append((a, b))
                start += m
        pieces = nxt
    return pieces


def _split_native(block, eff):
    """_split of the whole block array by the library (cmp_cluster_blocks): the same boxes in the same
    order."""
    import ctypes
    from . import capi
    b = np.ascontiguousarray(block, dtype=np.uint8)
    dim = b.ndim
    nmax = max(1, int(b.sum()))
    shape = (ctypes.c_int * dim)(*b.shape)
    out = (ctypes.c_int * (6 * nmax))()
    n = ctypes.c_int()
    rc = capi.load().cmp_cluster_blocks(b.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), dim, shape,
                                        float(eff), nmax, out, ctypes.byref(n))
    if rc != capi.CMP_OK:
        raise RuntimeError("cmp_cluster_blocks failed: %s" % capi.ERROR_NAMES.get(rc, rc))
    return [(list(out[6 * q:6 * q + dim]), list(out[6 * q + 3:6 * q + 3 + dim])) for q in range(n.value)]


def cluster(tags, blocking_factor=8, max_grid_size=64, grid_eff=0.7, n_error_buf=1, scale=1, n_cell=None,
            native=True):
    """Boxes [(lo, hi)] (cell indices, x first, inclusive) covering every tag of `tags`
    ([(k,) j, i]), aligned to the blocking factor, at most max_grid_size cells a side. scale /
    n_cell: the tags stand for blocks of `scale` cells of a grid n_cell ([(k,) j, i]) in size, and
    the boxes are returned in those cells. native: the split step in the library (the same boxes as
    the Python _split, which native=False runs)."""
    dim = tags.ndim
    t = _grow(tags, n_error_buf)
    block = _coarsen(t, blocking_factor)
    if native:
        raw = _split_native(block, grid_eff)
    else:
        raw = []
        _split(block, [0] * dim, [s - 1 for s in block.shape], grid_eff, raw)
    maxb = max(1, max_grid_size // blocking_factor)
    out = []
    f = blocking_factor * scale
    n_cell = tags.shape if n_cell is None else tuple(n_cell)
    for lo, hi in raw:
        for a, b in _chop(lo, hi, maxb):
            clo = [a[ax] * f for ax in range(dim)]
            chi = [min((b[ax] + 1) * f, n_cell[ax]) - 1 for ax in range(dim)]
            out.append((list(reversed(clo)), list(reversed(chi))))   # numpy axes -> x first
    return out


def refine(boxes, ratio):
    return [([l * ratio for l in lo], [(h + 1) * ratio - 1 for h in hi]) for lo, hi in boxes]


def tags_to_global(layout, tags, shape):
    """The one-byte-per-valid-cell tags of cmp_level_tag (box by box) as a global array."""
    out = np.zeros(shape, dtype=np.uint8)
    q = 0
    for lo, hi in layout.boxes:
        ext = [h - l + 1 for l, h in zip(lo, hi)]
        n = int(np.prod(ext))
        sl = tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(len(lo))))
        out[sl] = tags[q:q + n].reshape(tuple(reversed(ext)))
        q += n
    return out


def regrid(level, U, criteria, shape, ratio=2, blocking_factor=8, max_grid_size=64, grid_eff=0.7,
           n_error_buf=1, stream=None):
    """The next level's boxes (in its cells) from this level's ghost-filled state U (torch tensor
    on the GPU): GPU tagging with `criteria`, then cluster and refine."""
    import torch
    n = sum(int(np.prod([h - l + 1 for l, h in zip(lo, hi)])) for lo, hi in level.layout.boxes)
    tags = torch.zeros(n, dtype=torch.uint8, device=U.device)
    level.ErrorEst(U, criteria, tags, stream=stream)
    g = tags_to_global(level.layout, tags.cpu().numpy(), shape)
    return refine(cluster(g, blocking_factor, max_grid_size, grid_eff, n_error_buf), ratio), g


def nest(boxes, parent_boxes, shape, blocking_factor=8, margin=4):
    """Proper nesting: the parts of `boxes` (cells of a level, x first) that lie at least `margin`
    cells inside the union of `parent_boxes` (the same level's own boxes), kept in whole
    blocking-factor blocks and re-formed into boxes (block runs along x, stacked where the runs of
    consecutive block rows agree). A level clustered from a non-base level's tags is nested this
    way before it is refined, so that its ghost cells' coarse sources (the two-level fill) and the
    faces of the parent's own coarse-fine boundary stay clear of it."""
    dim = len(shape)
    cover = np.zeros(shape, dtype=bool)
    for lo, hi in parent_boxes:
        cover[tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(dim)))] = True
    inner = ~_grow(~cover, margin)
    bf = blocking_factor
    nbk = [-(-n // bf) for n in shape]
    keep = np.zeros(nbk, dtype=bool)
    for lo, hi in boxes:
        sl = tuple(slice(lo[d] // bf, hi[d] // bf + 1) for d in reversed(range(dim)))
        keep[sl] = True
    # a block stays when all its cells are inner
    pad = [(0, s * bf - n) for s, n in zip(nbk, shape)]
    t = np.pad(inner, pad)
    for ax in range(dim):
        t = t.reshape(t.shape[:ax] + (nbk[ax], bf) + t.shape[ax + 1:]).all(axis=ax + 1)
    keep &= t
    if dim != 2:
        # the kept blocks as boxes: Berger-Rigoutsos on them with efficiency 1 (every box entirely of
        # kept blocks), at most 64 cells a side
        return cluster(keep.astype(np.uint8), blocking_factor=1, max_grid_size=max(1, 64 // bf), grid_eff=1.0,
                       n_error_buf=0, scale=bf, n_cell=shape)
    out = []
    open_runs = {}          # (i0, i1) -> j0 of the run stack still growing
    for j in range(nbk[0] + 1):
        runs = set()
        if j < nbk[0]:
            i = 0
            while i < nbk[1]:
                if keep[j, i]:
                    i0 = i
                    while i < nbk[1] and keep[j, i] and i - i0 < 64 // bf:
                        i += 1
                    runs.add((i0, i - 1))
                else:
                    i += 1
        for r in list(open_runs):
            if r not in runs or j - open_runs[r] >= 64 // bf:
                j0 = open_runs.pop(r)
                out.append(([r[0] * bf, j0 * bf], [min((r[1] + 1) * bf, shape[1]) - 1, min(j * bf, shape[0]) - 1]))
        for r in runs:
            if r not in open_runs:
                open_runs[r] = j
    return out


def remake_level(params, geom, new_boxes, old_level, U_old, coarse, Uc, ratio=2, ng=3):
    """A level on new boxes after a regrid (RemakeLevel; MakeNewLevelFromCoarse when old_level is
    None): a new Level, its Layout and its state (every cell from the old level's boxes where they
    hold it, else interpolated from the coarse level, cmp_level_fill_from). Returns
    (level, layout, U)."""
    import torch
    from . import Layout, Level
    lay = Layout(params.dim, new_boxes, ng, params.nstate)
    lev = Level(params, geom, lay)
    U = torch.zeros(lay.total, dtype=torch.float64, device=Uc.device)
    lev.FillFrom(U, old_level, U_old, coarse, Uc, ratio)
    return lev, lay, U
