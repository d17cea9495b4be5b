"""A hierarchy advanced with regrids inside the run (Evolve's regrid every amr.regrid_int steps,
source/COMPAS.cpp:2141-2145, AmrCore::regrid -> ErrorEst / MakeNewGrids / RemakeLevel): between
groups of regrid_int steps (cmp_amr_advance, one step a replayed graph), every finer level is made
again from the GPU tags of the level below: the tags are clustered on the host (regrid.cluster),
nested inside that level's boxes, refined, and the new level takes its old data where its old boxes
hold it and the coarse level's conservative interpolation elsewhere (regrid.remake_level); then the
levels are averaged down and the hierarchy (cmp_amr, its step graph and, with the Phase-Field step
on, its Phase-Field buffers) is built again. With `tiles`, every level takes the tile autotuner's
choice before the hierarchy is built (cmp_level_tune_tiles: box shapes not swept yet are swept
then, so a run that pre-sweeps the shapes its regrids can make, `tile_shapes`, sweeps nothing
later), and the hierarchy its single work-list launch when the swept times favour it.
"""
import time

import numpy as np

from . import Amr, regrid


class RegridHierarchy:
    """levels / Us: the hierarchy's Levels and state tensors (coarsest first); geoms their
    geometries. criteria: cmp_level_tag criteria; blocking_factor / max_grid_size in the cells of the
    level being made; nest_margin in the cells of the level below it."""

    def __init__(self, params, geoms, levels, Us, ratio=2, criteria=None, blocking_factor=8, max_grid_size=64,
                 grid_eff=0.7, n_error_buf=1, nest_margin=4, pf=None, tiles=None, tile_shapes=()):
        import torch
        self.P, self.geoms, self.levels, self.Us = params, list(geoms), list(levels), list(Us)
        self.ratio, self.criteria = ratio, criteria
        self.bf, self.mgs, self.eff, self.nbuf, self.margin = blocking_factor, max_grid_size, grid_eff, n_error_buf, nest_margin
        self.pf = pf
        # tile autotuner: None / "fixed" off, "auto" by the swept times, "list" the work-list kernel
        self.tiles = tiles
        self.tile_shapes = list(tile_shapes)
        self.amr = None
        self.stats = dict(regrids=0, tag_and_cluster_s=0.0, remake_s=0.0, tune_s=0.0)
        self.updates = 0     # cell-updates of every step advanced (the cells of the hierarchy at that step)
        self.dt = torch.zeros(1, dtype=torch.float64, device=self.Us[0].device)
        self._build()

    def _build(self):
        import torch
        if self.amr is not None:
            self.amr.close()
        force = {"auto": 0, "list": 1}.get(self.tiles)
        if force is not None:
            # the remake's GPU work first, so that the tuner's time is its own
            torch.cuda.synchronize()
            t = time.perf_counter()
            for lv, U in zip(self.levels, self.Us):
                lv.tune_tiles(U, extra_shapes=self.tile_shapes, force=force)
            self.stats["tune_s"] += time.perf_counter() - t
        self.amr = Amr(self.levels, self.ratio)
        if force == 1:
            self.amr.tiles(1)
        if self.pf is not None:
            self.amr.pf_setup(self.pf)
        self.S1 = [torch.zeros_like(u) for u in self.Us]
        self.S2 = [torch.zeros_like(u) for u in self.Us]
        self.cmax = torch.zeros(len(self.levels), dtype=torch.float64, device=self.Us[0].device)

    def _shape(self, l):
        g = self.geoms[l]
        return tuple(int(g.domain_hi[d] - g.domain_lo[d] + 1) for d in reversed(range(self.P.dim)))

    def regrid(self):
        """Every finer level made again from the tags of the level below, coarsest first."""
        import torch
        r = self.ratio
        # the hierarchy object refers to the levels about to be replaced: it goes first
        if self.amr is not None:
            self.amr.close()
            self.amr = None
        for l in range(len(self.levels) - 1):
            t0 = time.perf_counter()
            lev, U = self.levels[l], self.Us[l]
            if l == 0:
                lev.FillPatch(U)
            else:
                lev.FillPatchTwoLevels(U, self.levels[l - 1], self.Us[l - 1], r)
            shape = self._shape(l)
            nb, _ = regrid.regrid(lev, U, self.criteria, shape, ratio=1,
                                  blocking_factor=max(1, self.bf // r), max_grid_size=max(1, self.mgs // r),
                                  grid_eff=self.eff, n_error_buf=self.nbuf)
            if l > 0:
                nb = regrid.nest(nb, lev.layout.boxes, shape, blocking_factor=max(1, self.bf // r), margin=self.margin)
            nb = regrid.refine(nb, r)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            self.stats["tag_and_cluster_s"] += t1 - t0
            if not nb:
                raise RuntimeError("regrid left level %d without boxes" % (l + 1))
            new, _, Unew = regrid.remake_level(self.P, self.geoms[l + 1], nb, self.levels[l + 1], self.Us[l + 1],
                                               lev, U, r, ng=self.levels[l + 1].layout.ng)
            err = new.check_errors()
            if err:
                raise RuntimeError("remake of level %d: device error %d" % (l + 1, err))
            self.levels[l + 1].close()
            self.levels[l + 1], self.Us[l + 1] = new, Unew
            torch.cuda.synchronize()
            self.stats["remake_s"] += time.perf_counter() - t1
        t2 = time.perf_counter()
        for l in range(len(self.levels) - 2, -1, -1):
            self.levels[l + 1].AverageDownTo(self.Us[l + 1], self.levels[l], self.Us[l], r)
        self._build()
        torch.cuda.synchronize()
        self.stats["remake_s"] += time.perf_counter() - t2
        self.stats["regrids"] += 1

    def advance(self, nsteps, cfl, regrid_int=0, limiter=1):
        """nsteps hierarchy steps (flow, then the Phase-Field step when on), a regrid after every
        regrid_int of them (0: none); the last group ends the call without a regrid."""
        done = 0
        while done < nsteps:
            k = nsteps - done if regrid_int <= 0 else min(regrid_int, nsteps - done)
            self.amr.advance(self.Us, self.S1, self.S2, k, cfl, self.dt, self.cmax, limiter=limiter)
            self.updates += k * sum(self.cells())
            done += k
            if regrid_int > 0 and done < nsteps:
                self.regrid()
        return done

    def cells(self):
        return [sum(int(np.prod([h - l + 1 for l, h in zip(lo, hi)])) for lo, hi in lv.layout.boxes)
                for lv in self.levels]

    def check_errors(self):
        return self.amr.check_errors()

    def close(self):
        if self.amr is not None:
            self.amr.close()
            self.amr = None
        for lv in self.levels:
            lv.close()
