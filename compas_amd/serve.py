"""The served stream: independent 2D two-phase cases arrive as a Poisson process, are admitted into a
batch of levels (cmp_batch_*) as case slots and device memory free up, and advance together, one
RK3 step of every admitted case per tick (one per-case ComputeDt launch and three stage launches
over the tiles of all of them). A case leaves the batch after its last step; its time to solution
is the time from its arrival to the end of that step.

This path has no counterpart in the reference, which runs one case per process
(source/main.cpp:21-86). The flow step is batched; with `pf` every tick also runs the Phase-Field
step of every admitted case (cmp_batch_pf_step: the cases concurrently, each on its own stream), as
each case's own run would after its flow step.
"""
import time

import numpy as np

from . import Batch, Layout, Level, cases


def _prepare(seeds, max_grid_size, grids):
    out = []
    for s in seeds:
        c = cases.served_case(s, max_grid_size=max_grid_size, grids=grids)
        P = c["params"]
        lay = Layout(P.dim, c["boxes"], 3, P.nstate)
        out.append(dict(case=c, layout=lay, U0=lay.pack(c["U"]), cells=int(np.prod(c["shape"]))))
    return out


def serve(rate, n_cases, steps=200, budget_bytes=16 << 30, max_cases=256, seed=0, max_grid_size=64,
          grids=None, stream=None, keep=None, pf=None, pf_threads=0):
    """Runs n_cases seeded cases arriving at `rate` cases/s through one batch whose three arenas take
    budget_bytes of device memory. Returns a dict: p50/p99/max time to solution (s), aggregate
    cell-updates/s over the run (first arrival to last completion), cells and cell-updates, ticks,
    mean cases per tick; `keep` (a set of case indices) also returns those cases' final states."""
    import torch
    rng = np.random.default_rng(seed)
    reqs = _prepare([seed * 100003 + i for i in range(n_cases)], max_grid_size, grids)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_cases))
    arrivals -= arrivals[0]
    like = reqs[0]["case"]["params"]
    batch = Batch(like, budget_bytes // (3 * 8), max_cases)
    if pf is not None:
        batch.pf_setup(pf)
    s = stream or torch.cuda.current_stream()
    pending = list(range(n_cases))
    active = {}                      # case id -> (request index, steps done)
    latency = np.zeros(n_cases)
    finals = {}
    ticks, occupancy, cell_updates = 0, 0, 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while pending or active:
        now = time.perf_counter() - t0
        # admit every arrived request, in order, while the batch has a slot and room for it
        while pending and arrivals[pending[0]] <= now:
            r = reqs[pending[0]]
            c = r["case"]
            cid = batch.admit(c["params"], c["geom"], r["layout"], r["U0"], c["cfl"], stream=s)
            if cid is None:
                break
            active[cid] = [pending.pop(0), 0]
        if not active:
            time.sleep(max(0.0, arrivals[pending[0]] - (time.perf_counter() - t0)))
            continue
        batch.step(stream=s)
        if pf is not None:
            batch.pf_step(pf_threads, stream=s)
        s.synchronize()
        ticks += 1
        occupancy += len(active)
        done = []
        for cid, st in active.items():
            st[1] += 1
            cell_updates += reqs[st[0]]["cells"]
            if st[1] == steps:
                done.append(cid)
        t_done = time.perf_counter() - t0
        for cid in done:
            i = active.pop(cid)[0]
            latency[i] = t_done - arrivals[i]
            if keep is not None and i in keep:
                finals[i] = batch.state(cid, stream=s)[0]
            err = batch.check_errors(cid, stream=s)
            if err:
                raise RuntimeError("served case %d ended with error %d" % (i, err))
            batch.retire(cid, stream=s)
    wall = time.perf_counter() - t0
    batch.close()
    return dict(rate=rate, cases=n_cases, steps=steps, p50_s=float(np.percentile(latency, 50)),
                p99_s=float(np.percentile(latency, 99)), max_s=float(latency.max()),
                cell_updates=int(cell_updates), wall_s=wall, cell_updates_per_s=cell_updates / wall,
                ticks=ticks, mean_cases_per_tick=occupancy / max(1, ticks), finals=finals)


def one_at_a_time(n_cases, steps=200, seed=0, max_grid_size=64, grids=None, pf=None):
    """The same cases run one after another, each alone through the single-level path (a captured
    step replayed, then the Phase-Field step with `pf`): cell-updates/s of that form."""
    import torch
    reqs = _prepare([seed * 100003 + i for i in range(n_cases)], max_grid_size, grids)
    dev = torch.device("cuda:0")
    total = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in reqs:
        c = r["case"]
        Un = torch.from_numpy(r["U0"]).to(dev)
        U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
        dt = torch.zeros(1, dtype=torch.float64, device=dev)
        cm = torch.zeros(1, dtype=torch.float64, device=dev)
        lev = Level(c["params"], c["geom"], r["layout"])
        if pf is not None:
            lev.pf_setup(pf)
        lev.advance(Un, U1, U2, steps, c["cfl"], dt, cm, use_graph=True)
        torch.cuda.synchronize()
        if lev.check_errors():
            raise RuntimeError("case ended with a device error")
        lev.close()
        total += r["cells"] * steps
    wall = time.perf_counter() - t0
    return dict(cases=n_cases, steps=steps, cell_updates=total, wall_s=wall, cell_updates_per_s=total / wall)


def plain_per_box(n_cases, steps=3, seed=0, max_grid_size=64, grids=None):
    """The same cases one at a time in the plain per-box form a direct port of the reference's box
    loop takes (compute_dUdt_FV per box: source/Class_Compressible_PhaseField.cpp:3473-3537, then
    the RK3 LinComb per box): per RK3 stage a ghost fill of the level, one cmp_rhs_box launch per box
    and the stage's combination of that box's valid cells; ComputeDt from the boxes' c_max on the
    host each step. Flow step only; cell-updates/s over `steps` steps of every case."""
    import torch
    from . import rhs_box
    reqs = _prepare([seed * 100003 + i for i in range(n_cases)], max_grid_size, grids)
    dev = torch.device("cuda:0")
    total = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in reqs:
        c, lay = r["case"], r["layout"]
        P, G = c["params"], c["geom"]
        N, ng = P.nstate, lay.ng
        lev = Level(P, G, lay)
        Un = torch.from_numpy(r["U0"]).to(dev)
        U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
        cm = torch.zeros(1, dtype=torch.float64, device=dev)
        views = []
        for b, (lo, hi) in enumerate(lay.boxes):
            off = lay.cboxes[b].offset
            shp = lay.fab_shape(b)
            size = int(np.prod(shp))
            nv = (hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
            sl = (slice(None), slice(ng, ng + nv[0]), slice(ng, ng + nv[1]))
            dU = torch.zeros((N,) + nv, dtype=torch.float64, device=dev)
            views.append((off, size, shp, sl, dU, (lo, hi)))
        dx = min(G.dx[0], G.dx[1])
        lev.FillPatch(Un)
        lev.CalculateInitialcmax(Un, cm)
        for _ in range(steps):
            dt = c["cfl"] * dx / float(cm.item())
            cm.zero_()
            for stage, (Us, Uo) in enumerate(((Un, U1), (U1, U2), (U2, Un))):
                lev.FillPatch(Us)
                a, b, w = ((0.0, 1.0, 1.0), (0.25, 0.75, 0.25), (2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0))[stage]
                for off, size, shp, sl, dU, box in views:
                    rhs_box(P, G, box, ng, Us[off:off + size], dU, cm if stage == 0 else None)
                    us = Us[off:off + size].view(shp)[sl]
                    un = Un[off:off + size].view(shp)[sl]
                    uo = Uo[off:off + size].view(shp)[sl]
                    if stage == 0:
                        uo.copy_(us + dt * dU)
                    else:
                        uo.copy_((a * us + b * un) + (w * dt) * dU)
            total += r["cells"]
        torch.cuda.synchronize()
        if lev.check_errors():
            raise RuntimeError("case ended with a device error")
        lev.close()
    wall = time.perf_counter() - t0
    return dict(cases=n_cases, steps=steps, cell_updates=total, wall_s=wall, cell_updates_per_s=total / wall)
