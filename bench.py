#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the five-equation WENO5-HLLC TVD-RK3 step on one GPU.

A step is one full time step of the level as the reference's Evolve loop takes it: device
ComputeDt and the flow's TVD-RK3 step (three ghost fills, three fused stage kernels) replayed from
a hipGraph, then, where the workload has interface regularization, the Phase-Field step
(mobility from max|u|, the ACDI mechanism sub-steps, Grad_Q, and Phase-Field RK3 sub-steps until
the flow step's dt is covered). The state stays resident in HBM. Prints one JSON line.

Workloads (--workload):
  water-air-droplet-3d    (default) 256^3 in 32^3 boxes, water droplet r = 0.1 (stiffened gas
                          gamma 4.4, p_inf 6e8 Pa, rho 1000) in air (gamma 1.4, rho 1.2), p = 1e5 Pa,
                          Mach 2.4 shock at x = 0.25, outflow BCs; WENO5 + THINC, HLLC, RK3 at
                          CFL 0.4; ACDI regularization (explicit RC, Phase-Field CFL 0.3,
                          eta = 0.5 dx)
  advected-interface-2d   256^2 periodic, two ideal gases (gamma 1.4/rho 1 in a circle r = 0.25,
                          gamma 1.6/rho 10 outside), u = v = p = 1, WENO5 + HLLC + RK3, CFL 0.5
  served-stream           seeded 2D shock-bubble cases (grids 128^2 .. 1024x256, Mach 1.1-3,
                          density ratio 0.1-10, gammas 1.1-1.7, radius 0.1-0.3) arriving as a
                          Poisson process at each rate of --rates, 200 RK3 steps each (WENO5 +
                          HLLC, CFL 0.4), advanced together in one batch of levels; p50/p99 time to
                          solution per rate and aggregate cell-updates/s, beside the same cases run
                          one at a time through the single-level path, and the plain per-box
                          form; every tick also runs each case's CAC Phase-Field step (the cases
                          concurrently; --no-pf: flow only).
  nphase-uneven-level     the AMR configuration's base level: 512^2 three-phase converging rings
                          (NPHASE = 3, WENO5 + HLLC, CFL 0.4) on a seeded regrid-like split into
                          644 boxes of 8..64 cells a side; the batched level step (one launch per
                          RK stage over the tiles of every box), and the right-hand side of the
                          level in one launch against the plain per-box form (one cmp_rhs_box
                          launch per box), both replayed from captured graphs. Flow step only.
  shock-bubble-amr        the shock-bubble configuration's flow step on its hierarchy: 1024x256 base
                          grid of [0,4]x[0,1] (boxes of 64) plus one level (ratio 2) from the
                          GPU tagging (|grad alpha| >= 0.1, |grad p| >= 0.1) clustered into boxes
                          of at most 64 (blocking factor 8); ambient gamma 1.4 rho 1 p 1, bubble
                          gamma 1.648 rho 0.138 r 0.25 at (1.5, 0.5), Mach 1.22 shock at x = 1;
                          WENO5B + HLLC, RK3 at CFL 0.4, levels stepped together without
                          subcycling (coarse faces under the fine level take its face means), a
                          regrid every --regrid-int steps inside the timed run (GPU tags, host
                          clustering, the fine level remade, graph rebuilt); each step followed by
                          the CAC Phase-Field step (composite implicit Grad_Q; --no-pf: flow only);
                          beside the uniform 2048x512 grid it stands for, stepped by the
                          single-level path with the same Phase-Field step.
  nphase-rings-amr        the AMR configuration's flow step on its hierarchy: the 512^2 three-phase
                          rings base level (uneven boxes of 8..64) plus three refined levels
                          (ratio 2), each clustered from the GPU tags of the level below (every
                          volume fraction's gradient >= 0.1, |grad p| >= 0.1; blocking factor 8,
                          boxes of at most 64, nested 4 cells inside their parent level), initial
                          data of each level from the same seeded rings at its resolution;
                          WENO5 + HLLC, RK3 at CFL 0.4, levels stepped together without
                          subcycling, a regrid every --regrid-int steps inside the timed run, the
                          CAC Phase-Field step after each step (--no-pf: flow only); beside the
                          uniform 4096^2 grid the finest level stands for.
Also reported: the dominant kernel's roofline position (measured with HIP events on the stream it
runs on) and the oracle (CPU restatement, OpenMP) timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector (= FP64 matrix) peak
PEAK_HBM_GBS = 8000.0


def workload(name):
    from compas_amd import cases, make_pf_params
    if name == "advected-interface-2d":
        c = cases.advected_interface_2d(n=256)
        c["config"] = {"workload": name, "grid": [256, 256], "boxes": len(c["boxes"]),
                       "scheme": "WENO5", "riemann": "HLLC", "integrator": "TVD-RK3", "cfl": 0.5,
                       "nstate": 6, "bc": "periodic", "phase_field": None}
        c["default_steps"] = 100
        c["pf"] = None
        return c
    if name == "water-air-droplet-3d":
        c = cases.water_air_droplet_3d(n=256, box=32)
        # PhaseField.* of the reference's water-air example inputs (ACDI, explicit RC, cfl 0.3,
        # Eta_Multiplier 0.5)
        c["pf"] = make_pf_params(mechanism="ACDI", explicit_rc=1, cfl=0.3, eta_multiplier=0.5)
        c["config"] = {"workload": name, "grid": [256, 256, 256], "boxes": len(c["boxes"]),
                       "box": [32, 32, 32], "scheme": "WENO5+THINC", "riemann": "HLLC",
                       "integrator": "TVD-RK3", "cfl": 0.4, "nstate": 7, "bc": "outflow",
                       "eos": "stiffened gas water / ideal air",
                       "phase_field": {"mechanism": "ACDI", "explicit_rc": 1, "flux": "UD mixture",
                                       "cfl": 0.3, "eta_multiplier": 0.5}}
        c["default_steps"] = 20
        return c
    raise SystemExit("unknown workload %r" % name)


def uneven_level(args):
    """The nphase-uneven-level workload: one JSON line (see the module docstring)."""
    import numpy as np
    import torch
    from compas_amd import Layout, Level, cases, rhs_box, roofline
    dev = torch.device("cuda:0")
    c = cases.nphase_rings_2d(512)
    P, G = c["params"], c["geom"]
    lay = Layout(2, c["boxes"], 3, P.nstate)
    cells = int(np.prod(c["shape"]))
    K = args.steps if args.steps is not None else 100
    W = args.warmup
    Un = torch.from_numpy(lay.pack(c["U"])).to(dev)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dt = torch.zeros(1, dtype=torch.float64, device=dev)
    cmax = torch.zeros(1, dtype=torch.float64, device=dev)
    lev = Level(P, G, lay)
    tune_s = tune_level(lev, Un, args)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        lev.advance(Un, U1, U2, W, c["cfl"], dt, cmax, use_graph=True, resume=False, stream=s)
        s.synchronize()
        t0 = time.perf_counter()
        lev.advance(Un, U1, U2, K, c["cfl"], dt, cmax, use_graph=True, resume=True, stream=s)
        s.synchronize()
        wall = time.perf_counter() - t0
    assert lev.check_errors(s) == 0
    # the right-hand side of the level: one launch over every tile, against one launch per box
    nv = [int(np.prod([h - l + 1 for l, h in zip(lo, hi)])) for lo, hi in c["boxes"]]
    dU_level = torch.zeros_like(Un)
    dU_boxes = torch.zeros(sum(nv) * P.nstate, dtype=torch.float64, device=dev)
    offs = np.concatenate([[0], np.cumsum(nv)]) * P.nstate
    graphs = {}
    for form in ("level", "per-box"):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cs = torch.cuda.current_stream()
            lev.FillPatch(Un, stream=cs)
            if form == "level":
                lev.compute_dUdt_FV(dU_level, Un, stream=cs)
            else:
                for b, box in enumerate(c["boxes"]):
                    o = int(lay.cboxes[b].offset)
                    rhs_box(P, G, box, 3, Un[o:], dU_boxes[int(offs[b]):], stream=cs)
        graphs[form] = g
    times = {}
    for form, g in graphs.items():
        g.replay()
        torch.cuda.synchronize()
        reps = 20
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        times[form] = e0.elapsed_time(e1) / reps
    # the two forms give the same right-hand side
    hl, hb = dU_level.cpu().numpy(), dU_boxes.cpu().numpy()
    same = all(np.array_equal(hb[int(offs[b]):int(offs[b + 1])],
                              lay.fab_view(hl, b)[:, 3:-3, 3:-3].ravel())
               for b in range(len(c["boxes"])))
    tiles_rec = tiles_record([lev], tune_s)
    lev.close()
    flops = roofline.flops_per_cell_stage(P) * cells
    rec = {
        "metric": "cell-updates/s (5Eq WENO5-HLLC RK3), % roofline; p50/p99 case latency",
        "value": cells * K / wall,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded three-phase converging rings)",
        "config": {"workload": "nphase-uneven-level", "grid": [512, 512], "boxes": len(c["boxes"]),
                   "box_sizes": "8..64 a side", "nphase": 3, "scheme": "WENO5", "riemann": "HLLC",
                   "integrator": "TVD-RK3", "cfl": 0.4, "phase_field": None},
        "rhs_ms": {"level_one_launch": times["level"], "per_box_launches": times["per-box"]},
        "batched_over_per_box": times["per-box"] / times["level"],
        "rhs_identical": bool(same),
        "tiles": tiles_rec,
        "roofline": {"bound": "valu", "kernel": "k_stage (fill + right-hand side of the level)",
                     "achieved": flops / (times["level"] * 1e-3) / 1e12, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": flops / (times["level"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                     "traffic": None, "flops_per_cell_stage": roofline.flops_per_cell_stage(P)},
    }
    if not args.no_cpu_baseline:
        cs = cases.nphase_rings_2d(128)
        cs["name"] = "the rings at 128^2 (uneven boxes)"
        c["name"] = "the rings base level at 512^2 (644 uneven boxes)"
        rec["cpu_baseline"] = cpu_baseline(level_sample(c, 1), level_sample(cs, 1))
    print(json.dumps(rec), flush=True)


def shock_bubble_amr(args):
    """The shock-bubble-amr workload: one JSON line (see the module docstring)."""
    import numpy as np
    import torch
    from compas_amd import Layout, Level, cases, regrid, roofline
    from compas_amd.amr_run import RegridHierarchy
    dev = torch.device("cuda:0")
    kw = dict(mach=1.22, rho_ratio=0.138, gamma_bubble=1.648, gamma_ambient=1.4, radius=0.25,
              max_grid_size=64, scheme="WENO5B", xb=1.5, xs=1.0)
    c0 = cases.shock_bubble_2d(1024, 256, **kw)
    P = c0["params"]
    l0 = Layout(2, c0["boxes"], 3, P.nstate)
    L0 = Level(P, c0["geom"], l0)
    U0 = torch.from_numpy(l0.pack(c0["U"])).to(dev)
    L0.FillPatch(U0)
    crit = [(5, "grad", 0.1), ("P", "grad", 0.1)]
    # (blocking factor 8 and max_grid_size 64 in the fine level's cells: 4 and 32 coarse cells)
    fine_boxes, tags = regrid.regrid(L0, U0, crit, c0["shape"], ratio=2, blocking_factor=4, max_grid_size=32)
    c1 = cases.shock_bubble_2d(2048, 512, **kw)
    l1 = Layout(2, fine_boxes, 3, P.nstate)
    L1 = Level(P, c1["geom"], l1)
    U1f = torch.from_numpy(l1.pack(c1["U"])).to(dev)
    L1.AverageDownTo(U1f, L0, U0, 2)
    pf = _amr_pf(args)
    init = [U0.cpu().numpy().copy(), U1f.cpu().numpy().copy()]   # the initial hierarchy, for the CPU baseline
    h = RegridHierarchy(P, [c0["geom"], c1["geom"]], [L0, L1], [U0, U1f], 2, criteria=crit, blocking_factor=8,
                        max_grid_size=64, pf=pf, tiles=None if args.tiles == "fixed" else args.tiles,
                        tile_shapes=_bf_shapes())
    cells = h.cells()
    K = args.steps if args.steps is not None else 50
    W = args.warmup
    h.advance(W, 0.4, args.regrid_int)
    torch.cuda.synchronize()
    u0 = h.updates
    st0 = dict(h.stats)
    t0 = time.perf_counter()
    h.advance(K, 0.4, args.regrid_int)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    updates = h.updates - u0
    assert h.check_errors() == 0
    regrid_rec = {"regrid_int": args.regrid_int, "regrids_in_timed_steps": h.stats["regrids"] - st0["regrids"],
                  "tag_and_cluster_s": h.stats["tag_and_cluster_s"] - st0["tag_and_cluster_s"],
                  "remake_and_rebuild_s": h.stats["remake_s"] - st0["remake_s"],
                  "tile_tuning_s": h.stats["tune_s"] - st0["tune_s"], "cells_per_level_at_end": h.cells()}
    tiles_rec = tiles_record(h.levels, st0["tune_s"], h.amr)
    pf_rec = _amr_pf_split(h.amr, h.Us, h.S1, h.S2, h.dt, pf)
    h.close()
    # the uniform fine grid the hierarchy stands for, single-level path (graph-replayed steps)
    lu = Layout(2, c1["boxes"], 3, P.nstate)
    Lu = Level(P, c1["geom"], lu)
    Uu = torch.from_numpy(lu.pack(c1["U"])).to(dev)
    tune_level(Lu, Uu, args)
    uniform_tiles = tiles_record([Lu])["levels"][0]
    if pf is not None:
        Lu.pf_setup(pf)
    Vu1, Vu2 = torch.zeros_like(Uu), torch.zeros_like(Uu)
    du = torch.zeros(1, dtype=torch.float64, device=dev)
    cu = torch.zeros(1, dtype=torch.float64, device=dev)
    Lu.advance(Uu, Vu1, Vu2, W, 0.4, du, cu, use_graph=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    Lu.advance(Uu, Vu1, Vu2, K, 0.4, du, cu, use_graph=True, resume=True)
    torch.cuda.synchronize()
    wall_u = time.perf_counter() - t1
    assert Lu.check_errors() == 0
    Lu.close()
    fps = roofline.flops_per_cell_stage(P)
    achieved = 3.0 * fps * updates / wall / 1e12
    rec = {
        "metric": "cell-updates/s (5Eq WENO5-HLLC RK3), % roofline; p50/p99 case latency",
        "value": updates / wall,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (shock-bubble initial state, hierarchy from the GPU tagging)",
        "config": {"workload": "shock-bubble-amr", "base_grid": [1024, 256], "levels": 2, "ratio": 2,
                   "fine_boxes": len(fine_boxes), "cells_per_level": cells,
                   "fine_coverage": cells[1] / (4.0 * cells[0]), "tagged_cells": int(tags.sum()),
                   "scheme": "WENO5B", "riemann": "HLLC", "integrator": "TVD-RK3 (levels stage-coupled)",
                   "cfl": 0.4, "phase_field": _amr_pf_config(pf)},
        "uniform_fine_grid": {"grid": [2048, 512], "ms_per_step": wall_u / K * 1e3,
                              "cell_updates_per_s": 2048 * 512 * K / wall_u},
        "amr_step_over_uniform_step": (wall / K) / (wall_u / K),
        "roofline": {"bound": "valu", "kernel": "whole hierarchy step (all kernels)", "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": None, "flops_per_cell_stage": fps},
    }
    rec["regrid"] = regrid_rec
    rec["tiles"] = tiles_rec
    rec["uniform_fine_grid"]["tiles"] = {k: uniform_tiles[k] for k in ("work_list_kernel", "waves_per_simd")}
    if pf_rec is not None:
        rec["phase_field"] = pf_rec
    if not args.no_cpu_baseline:
        cs = cases.shock_bubble_2d(256, 64, **kw)
        cs["name"] = "the same shock-bubble state at 256x64 (one level)"
        rec["cpu_baseline"] = cpu_baseline(
            amr_sample(P, [c0["geom"], c1["geom"]], [l0, l1], init, 1, 0.4, "the shock-bubble hierarchy"),
            level_sample(cs, 1))
        rec["cpu_baseline"]["note"] = ("flow steps only: the oracle's composite implicit Grad_Q is Jacobi-preconditioned "
                                       "CG, not the reference's MLMG, so its time says nothing about the reference")
    print(json.dumps(rec), flush=True)


def _tiles_force(args):
    return {"auto": 0, "list": 1, "fixed": -1}[args.tiles]


def _bf_shapes(bf=8, mgs=64):
    """Every box shape a regrid with blocking factor bf and max_grid_size mgs can make (2D)."""
    return [(i, j) for i in range(bf, mgs + 1, bf) for j in range(bf, mgs + 1, bf)]


def tiles_record(levels, tune_s=None, amr=None):
    """The tile choice of every level for the bench line: work-list kernel on, its occupancy target,
    how many boxes each kernel steps, and per box shape (the 8 holding the most cells) the chosen tile
    ("fixed" for a box the fixed-shape kernel keeps) and the swept times (ns per cell) of the fixed
    kernel and of that tile; the swept sums over the level's cells."""
    import numpy as np
    out = []
    for lv in levels:
        on, waves, ext = lv.tile_shapes()
        cells, chosen = {}, {}
        nfixed = 0
        for (lo, hi), e in zip(lv.layout.boxes, ext):
            n = tuple(h - l + 1 for l, h in zip(lo, hi))
            cells[n] = cells.get(n, 0) + int(np.prod(n))
            chosen[n] = tuple(e[:len(n)])
            nfixed += int(not on or e[0] == 0)
        t_fixed = t_chosen = 0.0
        top = {}
        swept = True
        for n in sorted(cells, key=lambda k: -cells[k]):
            t = lv.tile_tuning(n)
            if t is None:
                swept = False
                continue
            ns = t["fixed"]
            if on and chosen[n][0] > 0:
                w = 1 if waves == 3 else 0
                ns = min((c[1 + w] for c in t["cand"] if c[0][:len(n)] == chosen[n]), default=t["fixed"])
            t_fixed += cells[n] * t["fixed"]
            t_chosen += cells[n] * ns
            if len(top) < 8:
                tile = "x".join(map(str, chosen[n])) if on and chosen[n][0] > 0 else "fixed"
                top["x".join(map(str, n))] = {"tile": tile,
                                              "ns_per_cell_fixed": round(t["fixed"], 4),
                                              "ns_per_cell_tile": round(ns, 4)}
        out.append({"work_list_kernel": on, "waves_per_simd": waves if on else None,
                    "boxes_fixed_kernel": nfixed, "boxes_work_list": len(ext) - nfixed,
                    "box_shapes": len(cells), "swept": swept, "by_box_shape": top,
                    "swept_stage_ms_fixed": t_fixed * 1e-6 if swept else None,
                    "swept_stage_ms_chosen": t_chosen * 1e-6 if swept else None})
    rec = {"levels": out}
    if tune_s is not None:
        rec["tune_s"] = tune_s
    if amr is not None:
        on, waves = amr.tile_info()
        rec["hierarchy_one_launch"] = on
        rec["hierarchy_waves_per_simd"] = waves if on else None
    return rec


def tune_level(lev, U, args, extra=()):
    """The tile autotuner on a level before the timed steps; seconds it took."""
    import torch
    if args.tiles == "fixed":
        return 0.0
    t = time.perf_counter()
    lev.tune_tiles(U, extra_shapes=extra, force=_tiles_force(args))
    torch.cuda.synchronize()
    return time.perf_counter() - t


def _amr_pf(args):
    """The hierarchy workloads' interface regularization: CAC-Adv with the implicit RC (CAC requires
    it), Phase-Field CFL 0.3, eta = 0.5 finest cells; None with --no-pf."""
    if args.no_pf:
        return None
    from compas_amd import make_pf_params
    return make_pf_params(mechanism="CAC-Adv", explicit_rc=0, cfl=0.3, eta_multiplier=0.5)


def _amr_pf_config(pf, grad_q="composite implicit solve over the levels (CG, Jacobi + level-0 multigrid V-cycle)"):
    if pf is None:
        return None
    return {"mechanism": "CAC-Adv", "explicit_rc": 0, "flux": "UD mixture", "cfl": 0.3, "eta_multiplier": 0.5,
            "grad_q": grad_q}


def _amr_pf_split(amr, Un, S1, S2, dt, pf):
    """One more Phase-Field step of the hierarchy after the timed steps, timed alone (it is inside
    every timed step too): its wall time, sub-step counts and composite CG iterations."""
    if pf is None:
        return None
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    info = amr.PhaseField_Evolve(Un, S1, S2, dt)
    ms = (time.perf_counter() - t) * 1e3
    assert amr.check_errors() == 0
    return {"ms_per_pf_step": ms, "mechanism_substeps": info[3], "pf_rk3_substeps": info[5],
            "max_grad_q": info[4], "composite_cg_iterations": info[6]}


def nphase_rings_amr(args):
    """The nphase-rings-amr workload: one JSON line (see the module docstring)."""
    import numpy as np
    import torch
    from compas_amd import Layout, Level, cases, regrid, roofline
    from compas_amd.amr_run import RegridHierarchy
    dev = torch.device("cuda:0")
    crit = [(6, "grad", 0.1), (7, "grad", 0.1), (8, "grad", 0.1), ("P", "grad", 0.1)]
    c = cases.nphase_rings_2d(512)
    P = c["params"]
    levels, lays, Us, geoms = [], [], [], []
    lay = Layout(2, c["boxes"], 3, P.nstate)
    levels.append(Level(P, c["geom"], lay))
    lays.append(lay)
    geoms.append(c["geom"])
    Us.append(torch.from_numpy(lay.pack(c["U"])).to(dev))
    n = 512
    tagged = []
    for lev in range(3):
        U = Us[lev]
        if lev == 0:
            levels[0].FillPatch(U)
        else:
            levels[lev].FillPatchTwoLevels(U, levels[lev - 1], Us[lev - 1], 2)
        # boxes of the next level in this level's cells (blocking factor 8 and 64-cell boxes at the
        # next level: 4 and 32 here), nested inside this level's own boxes for the finer levels
        nb, tags = regrid.regrid(levels[lev], U, crit, (n, n), ratio=1, blocking_factor=4, max_grid_size=32)
        tagged.append(int(tags.sum()))
        if lev > 0:
            nb = regrid.nest(nb, lays[lev].boxes, (n, n), blocking_factor=4, margin=4)
        nb = regrid.refine(nb, 2)
        n *= 2
        cf = cases.nphase_rings_2d(n, boxes=nb)
        lay = Layout(2, nb, 3, P.nstate)
        levels.append(Level(P, cf["geom"], lay))
        lays.append(lay)
        geoms.append(cf["geom"])
        Us.append(torch.from_numpy(lay.pack(cf["U"])).to(dev))
        del cf
    for lev in range(len(levels) - 2, -1, -1):
        levels[lev + 1].AverageDownTo(Us[lev + 1], levels[lev], Us[lev], 2)
    init = [u.cpu().numpy().copy() for u in Us] if not args.no_cpu_baseline else None
    pf = _amr_pf(args)
    h = RegridHierarchy(P, geoms, levels, Us, 2, criteria=crit, blocking_factor=8, max_grid_size=64, nest_margin=4,
                        pf=pf, tiles=None if args.tiles == "fixed" else args.tiles, tile_shapes=_bf_shapes())
    cells = h.cells()
    K = args.steps if args.steps is not None else 50
    W = args.warmup
    h.advance(W, 0.4, args.regrid_int)
    torch.cuda.synchronize()
    u0 = h.updates
    st0 = dict(h.stats)
    t0 = time.perf_counter()
    h.advance(K, 0.4, args.regrid_int)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    updates = h.updates - u0
    assert h.check_errors() == 0
    regrid_rec = {"regrid_int": args.regrid_int, "regrids_in_timed_steps": h.stats["regrids"] - st0["regrids"],
                  "tag_and_cluster_s": h.stats["tag_and_cluster_s"] - st0["tag_and_cluster_s"],
                  "remake_and_rebuild_s": h.stats["remake_s"] - st0["remake_s"],
                  "tile_tuning_s": h.stats["tune_s"] - st0["tune_s"], "cells_per_level_at_end": h.cells()}
    tiles_rec = tiles_record(h.levels, st0["tune_s"], h.amr)
    pf_rec = _amr_pf_split(h.amr, h.Us, h.S1, h.S2, h.dt, pf)
    h.close()
    del Us
    torch.cuda.empty_cache()
    # the uniform finest grid, single-level path (graph-replayed steps)
    cu = cases.nphase_rings_2d(n, boxes=cases.split_domain((n, n), 64))
    lu = Layout(2, cu["boxes"], 3, P.nstate)
    Lu = Level(P, cu["geom"], lu)
    Uu = torch.from_numpy(lu.pack(cu["U"])).to(dev)
    del cu
    tune_level(Lu, Uu, args)
    uniform_tiles = tiles_record([Lu])["levels"][0]
    if pf is not None:
        Lu.pf_setup(pf)
    Vu1, Vu2 = torch.zeros_like(Uu), torch.zeros_like(Uu)
    du = torch.zeros(1, dtype=torch.float64, device=dev)
    cmu = torch.zeros(1, dtype=torch.float64, device=dev)
    Ku = max(2, K // 5)
    Lu.advance(Uu, Vu1, Vu2, 2, 0.4, du, cmu, use_graph=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    Lu.advance(Uu, Vu1, Vu2, Ku, 0.4, du, cmu, use_graph=True, resume=True)
    torch.cuda.synchronize()
    wall_u = time.perf_counter() - t1
    assert Lu.check_errors() == 0
    Lu.close()
    fps = roofline.flops_per_cell_stage(P)
    achieved = 3.0 * fps * updates / wall / 1e12
    rec = {
        "metric": "cell-updates/s (5Eq WENO5-HLLC RK3), % roofline; p50/p99 case latency",
        "value": updates / wall,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded three-phase converging rings, hierarchy from the GPU tagging)",
        "config": {"workload": "nphase-rings-amr", "base_grid": [512, 512], "levels": len(levels), "ratio": 2,
                   "boxes_per_level": [la.nbox for la in lays], "cells_per_level": cells,
                   "tagged_cells": tagged, "nphase": 3, "scheme": "WENO5", "riemann": "HLLC",
                   "integrator": "TVD-RK3 (levels stage-coupled)", "cfl": 0.4,
                   "phase_field": _amr_pf_config(pf)},
        "uniform_finest_grid": {"grid": [n, n], "ms_per_step": wall_u / Ku * 1e3,
                                "cell_updates_per_s": n * n * Ku / wall_u},
        "amr_step_over_uniform_step": (wall / K) / (wall_u / Ku),
        "roofline": {"bound": "valu", "kernel": "whole hierarchy step (all kernels)", "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": None, "flops_per_cell_stage": fps},
    }
    rec["regrid"] = regrid_rec
    rec["tiles"] = tiles_rec
    rec["uniform_finest_grid"]["tiles"] = {k: uniform_tiles[k] for k in ("work_list_kernel", "waves_per_simd")}
    if pf_rec is not None:
        rec["phase_field"] = pf_rec
    if init is not None:
        cs = cases.nphase_rings_2d(128)
        cs["name"] = "the rings at 128^2 (one level, uneven boxes)"
        rec["cpu_baseline"] = cpu_baseline(amr_sample(P, geoms, lays, init, 1, 0.4, "the rings hierarchy"),
                                           level_sample(cs, 1))
        rec["cpu_baseline"]["note"] = ("flow steps only: the oracle's composite implicit Grad_Q is Jacobi-preconditioned "
                                       "CG, not the reference's MLMG, so its time says nothing about the reference")
    print(json.dumps(rec), flush=True)


def served(args):
    """The served-stream workload: one JSON line (see the module docstring)."""
    import numpy as np
    import oracle
    from compas_amd import Layout, cases, roofline, serve
    rates = [float(r) for r in args.rates.split(",") if r]
    W = args.warmup
    pf = _amr_pf(args)   # CAC-Adv, implicit RC (the served configuration's regularization)
    # warm-up: a few small cases through the same loop (kernels loaded, allocator primed)
    serve.serve(max(rates), max(2, W), steps=5, seed=999, pf=pf, pf_threads=args.pf_threads)
    runs = []
    for r in rates:
        res = serve.serve(r, args.cases, steps=args.case_steps, seed=1, pf=pf, pf_threads=args.pf_threads)
        res.pop("finals")
        runs.append(res)
    alone = serve.one_at_a_time(args.cases, steps=args.case_steps, seed=1, pf=pf)
    # the plain per-box GPU form (flow step only, against the batched flow step below)
    plain = serve.plain_per_box(args.cases, steps=3, seed=1)
    flow = serve.serve(rates[-1], args.cases, steps=args.case_steps, seed=1)
    flow.pop("finals")
    top = runs[-1]
    P = cases.served_case(0)["params"]
    achieved = roofline.flops_per_cell_stage(P) * 3 * top["cell_updates_per_s"] / 1e12
    rec = {
        "metric": "cell-updates/s (5Eq WENO5-HLLC RK3), % roofline; p50/p99 case latency",
        "value": top["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.case_steps,
        "warmup": W,
        "ms_per_step": top["wall_s"] / max(1, top["ticks"]) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded shock-bubble cases)",
        "config": {"workload": "served-stream", "cases_per_rate": args.cases, "steps_per_case": args.case_steps,
                   "grids": [list(g) for g in cases.SERVED_GRIDS], "scheme": "WENO5", "riemann": "HLLC",
                   "integrator": "TVD-RK3", "cfl": 0.4, "max_grid_size": 64, "phase_field": _amr_pf_config(
                       pf, "implicit solve per case (CG preconditioned by the case's multigrid V-cycle); the "
                           "cases' Phase-Field steps concurrent, each on its own stream"),
                   "value_at_rate": rates[-1]},
        "latency": [{k: v for k, v in run.items()} for run in runs],
        "one_at_a_time": alone,
        "batched_over_one_at_a_time": top["cell_updates_per_s"] / alone["cell_updates_per_s"],
        "plain_per_box": dict(plain, form="one case at a time: per RK3 stage a level ghost fill, one cmp_rhs_box "
                                          "launch per box and the stage combination of each box; flow step only"),
        "batched_flow_only": {k: flow[k] for k in ("p50_s", "p99_s", "cell_updates_per_s", "wall_s", "ticks")},
        "batched_flow_over_plain_per_box": flow["cell_updates_per_s"] / plain["cell_updates_per_s"],
        "roofline": {"bound": "valu", "kernel": "k_stage_batch (whole request loop: host work and gaps included)",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "flops_per_cell_stage": roofline.flops_per_cell_stage(P)},
    }
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = served_cpu_baseline(args.case_steps)
    print(json.dumps(rec), flush=True)


def served_cpu_baseline(case_steps, ncases=4, sample_steps=5):
    """The served stream's cases on the CPU one at a time at all cores (as a CPU server would run
    them): per case the rate of `sample_steps` steps (1 warm-up step first) and its time to
    solution for `case_steps` steps extrapolated from it; p50 / p99 over `ncases` cases of the
    stream's seed sequence, aggregate cell-updates/s; plus the cpu_baseline record of the first case
    (all cores and 1 thread, median of 5)."""
    import numpy as np
    import oracle
    from compas_amd import cases
    oracle.set_num_threads(_threads())
    tts, upd, tsum = [], 0, 0.0
    for k in range(ncases):
        c = cases.served_case((k + 1) * 100003)
        smp = level_sample(c, sample_steps)
        el, _ = _time_runs(smp["reset"], smp["run"], 1)
        tts.append(el / sample_steps * case_steps)
        upd += smp["updates"]
        tsum += el
    c0 = cases.served_case(100003)
    c0["name"] = "served case 1 (%dx%d)" % (c0["shape"][1], c0["shape"][0])
    rec = cpu_baseline(level_sample(c0, 5), level_sample(c0, 1), native=False)
    rec["served_one_at_a_time"] = {
        "cases": ncases, "steps_per_case": case_steps, "cores": _threads(),
        "p50_s": float(np.percentile(tts, 50)), "p99_s": float(np.percentile(tts, 99)),
        "cell_updates_per_s": upd / tsum,
        "note": "time to solution of %d steps extrapolated from %d timed steps per case" % (case_steps, sample_steps)}
    return rec


def _host():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def _threads():
    """All host cores this process may use (OMP_NUM_THREADS caps it where the host sets it)."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def _time_runs(reset, run, runs):
    """1 warm-up run, then `runs` timed runs; reset() (untimed) before each. Returns the median time
    and the list."""
    import statistics
    reset()
    run()
    ts = []
    for _ in range(runs):
        reset()
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), ts


def level_sample(c, steps, pf=None):
    """A CPU sample: `steps` full time steps of a single level (RK3 flow step with ComputeDt, then
    the Phase-Field step when pf is given) from the case's initial state."""
    import numpy as np
    import oracle
    from compas_amd import Layout
    P, G = c["params"], c["geom"]
    lay = Layout(P.dim, c["boxes"], c.get("ng", 3), P.nstate)
    buf0 = lay.pack(c["U"])
    st = {}

    def reset():
        st["b"] = buf0.copy()

    def run():
        if pf is None:
            oracle.level_advance(P, G, lay, st["b"], steps, c["cfl"])
        else:
            oracle.level_advance_pf(P, pf, G, lay, st["b"], steps, c["cfl"])
    cells = sum(int(np.prod([h - l + 1 for l, h in zip(lo, hi)])) for lo, hi in c["boxes"])
    return dict(reset=reset, run=run, updates=cells * steps,
                what="%d %s of %s (%d cells)" % (steps, "flow + Phase-Field steps" if pf is not None else "RK3 steps",
                                                 c.get("name", "the case"), cells))


def amr_sample(P, geoms, lays, bufs, steps, cfl, name):
    """A CPU sample: `steps` flow steps of a hierarchy stepped together without subcycling."""
    import numpy as np
    import oracle
    st = {}

    def reset():
        st["b"] = [b.copy() for b in bufs]

    def run():
        oracle.amr_advance(P, geoms, lays, st["b"], steps, cfl, 2)
    cells = sum(sum(int(np.prod([h - l + 1 for l, h in zip(lo, hi)])) for lo, hi in la.boxes) for la in lays)
    return dict(reset=reset, run=run, updates=cells * steps,
                what="%d hierarchy flow steps of %s (%d levels, %d cells)" % (steps, name, len(lays), cells))


def cpu_baseline(sample_all, sample_one, runs=5, native=True):
    """The oracle (reference-structure CPU restatement, gcc -O3 -ffp-contract=off -fopenmp, no
    fast-math) timed on this host: 1 warm-up run, then the median of `runs` runs of sample_all at all
    host cores and of sample_one at 1 thread; with native, sample_all again with a -march=native
    build of the same source compiled here. Samples come from level_sample / amr_sample."""
    import oracle
    import tempfile
    from oracle import build as obuild
    nall = _threads()
    oracle.set_num_threads(nall)
    med, ts = _time_runs(sample_all["reset"], sample_all["run"], runs)
    rec = {"value": sample_all["updates"] / med, "unit": "cell-updates/s", "cores": nall, "kind": "port",
           "sample": sample_all["what"] + ", median of %d runs after 1 warm-up" % runs,
           "runs_s": ts, "build": "gcc " + " ".join(obuild.CFLAGS)}
    rec.update(_host())
    oracle.set_num_threads(1)
    med1, ts1 = _time_runs(sample_one["reset"], sample_one["run"], runs)
    rec["one_thread"] = {"value": sample_one["updates"] / med1, "unit": "cell-updates/s", "cores": 1,
                         "sample": sample_one["what"] + ", median of %d runs after 1 warm-up" % runs, "runs_s": ts1}
    oracle.set_num_threads(nall)
    if native:
        try:
            with tempfile.TemporaryDirectory() as d:
                path = obuild.build_variant(os.path.join(d, "libcpu_ref_native.so"), ["-march=native"])
                oracle.use_library(path)
                oracle.set_num_threads(nall)
                medn, tsn = _time_runs(sample_all["reset"], sample_all["run"], runs)
                rec["march_native"] = {"value": sample_all["updates"] / medn, "unit": "cell-updates/s",
                                       "cores": nall, "runs_s": tsn,
                                       "build": "gcc " + " ".join(obuild.CFLAGS + ["-march=native"])}
        except Exception as e:   # a host without a compiler: the variant is reported as missing
            rec["march_native"] = {"error": str(e)[:200]}
        finally:
            oracle.use_library(None)
    return rec


def measure_traffic(workload_name, kernel_prefix, calib_mib=512, extra=()):
    """HBM bytes per launch of the dominant kernel from two rocprofv3 counter runs of this script
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads for
    an in-place scaling of calib_mib MiB (reads and writes each byte once; a kernel, where torch's
    device-to-device copy may run on a DMA engine the counters do not see) run first in the same
    process. The kernel's last
    dispatch is the stage-timing launch the roofline's flop rate is measured on. None when
    rocprofv3 is missing or a run fails."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    rp = shutil.which("rocprofv3")
    if rp is None:
        return None
    child = [sys.executable, os.path.abspath(__file__), "--workload", workload_name, "--steps", "1",
             "--warmup", "1", "--no-cpu-baseline", "--kernel-reps", "2",
             "--copy-calibration-mib", str(calib_mib), "--no-traffic"] + list(extra)
    env = dict(os.environ, TMPDIR="/tmp")
    got = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="cmp_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "300", rp, "--pmc", ctr, "--output-format", "csv", "-d", d,
               "-o", "run", "--"] + child
        r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, env=env)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not files:
            return None
        kern, copy = [], []
        for row in csv.DictReader(open(files[0])):
            if row["Counter_Name"] != ctr:
                continue
            name = row["Kernel_Name"]
            if name.startswith(kernel_prefix):
                kern.append((int(row["Dispatch_Id"]), float(row["Counter_Value"])))
            elif "MulFunctor" in name:
                copy.append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if not kern or not copy:
            return None
        scale = calib_mib * 1024.0 / max(copy)        # the calibration pass is the largest product
        got[ctr] = (max(kern)[1] * scale, scale)
    fetch, fs = got["FETCH_SIZE"]
    write, ws = got["WRITE_SIZE"]
    return {"bytes": (fetch + write) * 1024.0, "fetch_bytes": fetch * 1024.0, "write_bytes": write * 1024.0,
            "scale_fetch": fs, "scale_write": ws}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="water-air-droplet-3d")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="scale this many MiB in place first: the reference kernel of known traffic "
                         "against which HBM counters (FETCH_SIZE, WRITE_SIZE) are scaled")
    ap.add_argument("--rates", default="5,50,200",
                    help="served-stream: arrival rates (cases/s) to run, comma separated")
    ap.add_argument("--cases", type=int, default=48, help="served-stream: cases per rate")
    ap.add_argument("--case-steps", type=int, default=200, help="served-stream: RK3 steps per case")
    ap.add_argument("--no-pf", action="store_true",
                    help="shock-bubble-amr, nphase-rings-amr, served-stream: the flow step alone, without the CAC Phase-Field step")
    ap.add_argument("--pf-threads", type=int, default=0,
                    help="served-stream: host workers driving the cases' Phase-Field steps (0: the library's default)")
    ap.add_argument("--regrid-int", type=int, default=10,
                    help="shock-bubble-amr, nphase-rings-amr: regrid every this many steps inside the timed run "
                         "(amr.regrid_int; 0: the hierarchy fixed)")
    ap.add_argument("--tiles", choices=("auto", "fixed", "list"), default="auto",
                    help="the tile autotuner before the timed steps: auto (the swept times choose between the "
                         "fixed-shape kernel and the persistent work-list kernel with per-box-shape tiles), "
                         "fixed, or list (the work-list kernel with the swept shapes)")
    ap.add_argument("--tile-cache", default=None,
                    help="load the tile autotuner's cache from this file first (cmp_tile_cache_load)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter runs that measure the dominant kernel's HBM traffic")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (every box of a "
              "level runs on one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from compas_amd import Layout, Level, roofline

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.tile_cache and os.path.exists(args.tile_cache):
        from compas_amd import tile_cache_load
        tile_cache_load(args.tile_cache)
    if args.workload == "served-stream":
        served(args)
        return
    if args.workload == "nphase-uneven-level":
        uneven_level(args)
        return
    if args.workload == "shock-bubble-amr":
        shock_bubble_amr(args)
        return
    if args.workload == "nphase-rings-amr":
        nphase_rings_amr(args)
        return
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        src.mul_(1.0)
        torch.cuda.synchronize()
        del src
    c = workload(args.workload)
    K = args.steps if args.steps is not None else c["default_steps"]
    W = args.warmup
    P, G = c["params"], c["geom"]
    lay = Layout(P.dim, c["boxes"], 3, P.nstate)
    buf = lay.pack(c["U"])
    Un = torch.from_numpy(buf).to(dev)
    U1, U2 = torch.zeros_like(Un), torch.zeros_like(Un)
    dt = torch.zeros(1, dtype=torch.float64, device=dev)
    cmax = torch.zeros(1, dtype=torch.float64, device=dev)
    lev = Level(P, G, lay)
    tune_s = tune_level(lev, Un, args)
    tiles_on, tiles_waves, _ = lev.tile_shapes()
    if c["pf"] is not None:
        lev.pf_setup(c["pf"])
    stream = torch.cuda.Stream()
    cells = int(np.prod(c["shape"]))

    with torch.cuda.stream(stream):
        lev.advance(Un, U1, U2, W, c["cfl"], dt, cmax, use_graph=True, resume=False, stream=stream)
        stream.synchronize()
        assert lev.check_errors(stream) == 0
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(stream)
        lev.advance(Un, U1, U2, K, c["cfl"], dt, cmax, use_graph=True, resume=True, stream=stream)
        ev1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        err = lev.check_errors(stream)
        assert err == 0, "device error word %d" % err
        wall = t1 - t0
        ev_ms = ev0.elapsed_time(ev1)

        # dominant kernel: the fused stage kernel as the step runs it (an RK3 step is its three
        # launches and nothing else; the last launch is the stage-2 form: reads U^(2) and U^n,
        # writes U^(n+1))
        k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k0.record(stream)
        for _ in range(args.kernel_reps):
            lev.ThirdOrderSSPRK(Un, U1, U2, dt, cmax, stream=stream)
        k1.record(stream)
        stream.synchronize()
        stage_ms = k0.elapsed_time(k1) / (3 * args.kernel_reps)
        assert lev.check_errors(stream) == 0

        pf_rec = None
        if c["pf"] is not None:
            # the Phase-Field step alone (host-synchronised: it reads max|u| and max|Grad_Q|)
            t2 = time.perf_counter()
            info = lev.PhaseField_Evolve(Un, U1, U2, dt, stream=stream)
            stream.synchronize()
            pf_ms = (time.perf_counter() - t2) * 1e3
            # its stage kernel: the Phase-Field right-hand side with the step's Grad_Q
            gq = torch.zeros(lev.faces.total, dtype=torch.float64, device=dev)
            lev.PhaseField_GradQPF(Un, float(dt.item()), gq, stream=stream)
            lev.FillPatch(U1, stream=stream)
            k0.record(stream)
            for _ in range(args.kernel_reps):
                lev.PhaseField_compute_dUdt_FV(U2, U1, gq, stream=stream)
            k1.record(stream)
            stream.synchronize()
            assert lev.check_errors(stream) == 0
            pf_rec = {"ms_per_pf_step": pf_ms, "mechanism_substeps": info[3],
                      "pf_rk3_substeps": info[5], "max_grad_q": info[4], "mobility": info[1],
                      "pf_rhs_kernel_ms": k0.elapsed_time(k1) / args.kernel_reps}

    flops_stage = roofline.flops_per_cell_stage(P) * cells
    achieved = flops_stage / (stage_ms * 1e-3) / 1e12
    bytes_update = roofline.bytes_per_cell_update(P)
    rec = {
        "metric": "cell-updates/s (5Eq WENO5-HLLC RK3), % roofline; p50/p99 case latency",
        "value": cells * K / wall,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded initial state of the named case)",
        "config": c["config"],
        "roofline": {
            "bound": "valu",
            "kernel": ("k_stage_list (persistent work-list, the autotuner's tile per box shape, %d waves/SIMD; "
                       % tiles_waves if tiles_on else "k_stage (")
                      + "fused %s + %s + divergence + RK3 stage)" % (c["config"]["scheme"], c["config"]["riemann"]),
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": None,
            "kernel_ms": stage_ms,
            "flops_per_cell_stage": roofline.flops_per_cell_stage(P),
            "hbm_GBps_algorithmic": bytes_update / 3.0 * cells / (stage_ms * 1e-3) / 1e9,
        },
        "event_ms_per_step": ev_ms / K,
    }
    rec["tiles"] = tiles_record([lev], tune_s)
    if pf_rec is not None:
        rec["phase_field"] = pf_rec
    if not args.no_traffic:
        scheme = {(1, 5): 5, (1, 6): 6, (0, 0): 0}.get((P.id_weno, P.id_limiter), 1)
        targs = "%d, %d, %d, %s, %d" % (P.dim, scheme, P.id_riemann, "true" if P.id_thinc else "false", P.nphase)
        prefix = ("void cmp::k_stage_list<%s, %d>" % (targs, tiles_waves) if tiles_on
                  else "void cmp::k_stage<%s>" % targs)
        # the counter runs take this run's tile choice from its cache (no sweep of their own)
        import tempfile
        from compas_amd import tile_cache_save
        cache = os.path.join(tempfile.mkdtemp(prefix="cmp_tiles_", dir="/tmp"), "tiles.txt")
        tile_cache_save(cache)
        tr = measure_traffic(args.workload, prefix,
                             extra=["--tiles", "list" if tiles_on else "fixed", "--tile-cache", cache])
        if tr is not None:
            rec["roofline"]["traffic"] = tr["bytes"]
            rec["roofline"]["traffic_detail"] = {
                "fetch_bytes": tr["fetch_bytes"], "write_bytes": tr["write_bytes"],
                "scale_fetch": tr["scale_fetch"], "scale_write": tr["scale_write"],
                "algorithmic_bytes": 3 * P.nstate * 8 * cells,
                "launch": "stage 2 form (reads U^(2) and U^n, writes U^(n+1); cells beyond a box read from their source boxes)"}
    if not args.no_cpu_baseline:
        from compas_amd import cases
        if args.workload == "water-air-droplet-3d":
            # the same case and boxes at 128^3 (all cores) and 64^3 (1 thread): the per-cell work of
            # the flow and Phase-Field steps, bounded to seconds
            ca = cases.water_air_droplet_3d(n=128, box=32)
            ca["name"] = "the water-air droplet at 128^3 in 32^3 boxes"
            c1 = cases.water_air_droplet_3d(n=64, box=32)
            c1["name"] = "the water-air droplet at 64^3 in 32^3 boxes"
            rec["cpu_baseline"] = cpu_baseline(level_sample(ca, 1, c["pf"]), level_sample(c1, 1, c["pf"]))
        else:
            ca = dict(c, name="the advected interface at 256^2")
            rec["cpu_baseline"] = cpu_baseline(level_sample(ca, 10), level_sample(ca, 2))
    lev.close()
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
