"""The reference's inputs files (AMReX ParmParse syntax) as this library's parameters.

`read(path)` / `parse(text)` give the key -> values table; `case_parameters(table, dim, nphase)` maps
the keys this library's path uses onto cmp_params, cmp_geometry and cmp_pf_params, with the
reference's defaults where a key is absent:

- FiniteVolume.Scheme / ID_Bound / Quad / THINC.* (include/Operators_FV/FiniteVolume_Parameter.H:50-205)
- Physics.source_term (include/Physics_Compressible5Eq/Physics_Parameter_Compressible5Eq.H:30-36),
  Physics.RiemannSolver (source/Class_Compressible_PhaseField.cpp:91-111)
- prob.EOS.EOS, gamma_k / pinf_k / b_k / q_k, or the per-phase lists gamma / pinf / b / q of the
  N-phase model (exec/*/prob/Parm.H; coefficients source/Parm.cpp:122-170, 230-262)
- geometry.prob_lo / prob_hi / is_periodic, amr.n_cell / max_grid_size / max_level, and with
  max_level > 0 amr.ref_ratio / regrid_int / n_error_buf / grid_eff / blocking_factor and
  run.do_subcycle (the hierarchy runs without subcycling), run.lo_bc / hi_bc / cfl /
  timestep_change_limiter / TimeIntegrator, max_step, stop_time (source/COMPAS.cpp:1186-1546)
- PhaseField.* (include/Physics_PhaseField/PhaseField_Parameter.H:29-140)

Settings the reference accepts but this library does not run are listed in "unsupported" (never
silently dropped); settings it does not use (output, plotting) are ignored.
"""
import shlex

from . import eos_coefficients, make_geometry, make_params, make_pf_params

# FiniteVolume.Scheme -> (id_weno, id_limiter, id_poly, ghost cells) (FiniteVolume_Parameter.H:58-147)
FV_SCHEMES = {
    "GODUNOV": (0, 0, 0, 1), "MUSCL": (0, 1, 1, 2), "MUSCL-Minmod": (0, 1, 1, 2),
    "MUSCL-vanLeer": (0, 2, 1, 2), "MUSCL-Superbee": (0, 3, 1, 2), "MUSCL-MC": (0, 4, 1, 2),
    "WENO3": (1, 3, 2, 2), "WENO3B": (1, 4, 2, 2), "WENO5": (1, 5, 4, 3), "WENO5B": (1, 6, 4, 3),
}
RIEMANN = {"LLF": 0, "HLL": 1, "HLLC": 2}
# run.TimeIntegrator (source/COMPAS.cpp:1306-1311) -> cmp_level_rk_step's integrator
TIME_INTEGRATORS = {"ForwardEuler": 1, "TVD-RK2": 2, "TVD-RK3": 3, "RK4": 4}


def parse(text):
    """ParmParse's `key = value ...` lines: '#' starts a comment, values split like a shell
    (quoted strings are one value), a later definition of a key replaces an earlier one."""
    table = {}
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line or "=" not in line:
            continue
        key, val = line.split("=", 1)
        table[key.strip()] = shlex.split(val.strip())
    return table


def read(path):
    with open(path) as f:
        return parse(f.read())


def _get(t, key, default, cast=float):
    v = t.get(key)
    return cast(v[0]) if v else default


def _arr(t, key, default, cast=float):
    v = t.get(key)
    return [cast(x) for x in v] if v else list(default)


def case_parameters(t, dim, nphase=0):
    """dict(params, geom, pf (None unless PhaseField.phase_field = 1), run, unsupported)."""
    unsupported = []
    # finite volume
    scheme = _get(t, "FiniteVolume.Scheme", "MUSCL", str)
    if scheme not in FV_SCHEMES:
        raise ValueError("FiniteVolume.Scheme = %r is not a scheme of the reference" % scheme)
    w, lim, poly, ng = FV_SCHEMES[scheme]
    bound = _get(t, "FiniteVolume.ID_Bound", 0, int)
    if scheme.startswith("MUSCL") and bound == 1:
        poly = 2
    if scheme in ("WENO3", "WENO3B"):
        unsupported.append("FiniteVolume.Scheme = %s" % scheme)
    if bound:
        unsupported.append("FiniteVolume.ID_Bound = 1")
    if _get(t, "FiniteVolume.Quad", 0, int):
        unsupported.append("FiniteVolume.Quad = 1")
    thinc = _get(t, "FiniteVolume.THINC", 0, int)
    if thinc:
        ng = max(ng, 2)
        if w == 0:
            unsupported.append("FiniteVolume.THINC = 1 on %s (THINC is built on WENO5 / WENO5B)" % scheme)
        for opt in ("Multi", "FaceState", "PhaseDensity"):
            if _get(t, "FiniteVolume.THINC." + opt, 0, int):
                unsupported.append("FiniteVolume.THINC.%s = 1" % opt)
    # physics
    source_term = _get(t, "Physics.source_term", 0, int)
    riemann = _get(t, "Physics.RiemannSolver", "HLLC", str)
    if riemann not in RIEMANN:
        raise ValueError("Physics.RiemannSolver = %r is not a solver of the reference" % riemann)
    if riemann == "LLF":
        unsupported.append("Physics.RiemannSolver = LLF")
    # equation of state
    eos = _get(t, "prob.EOS.EOS", 0, int)
    nph = nphase if nphase else 2
    if nphase:
        gamma = _arr(t, "prob.EOS.gamma", [1.4] * nph)
        pinf = _arr(t, "prob.EOS.pinf", [0.0] * nph)
        b = _arr(t, "prob.EOS.b", [0.0] * nph)
        q = _arr(t, "prob.EOS.q", [0.0] * nph)
    else:
        gamma = [_get(t, "prob.EOS.gamma_%d" % k, 1.4) for k in (1, 2)]
        pinf = [_get(t, "prob.EOS.pinf_%d" % k, 0.0) for k in (1, 2)]
        b = [_get(t, "prob.EOS.b_%d" % k, 0.0) for k in (1, 2)]
        q = [_get(t, "prob.EOS.q_%d" % k, 0.0) for k in (1, 2)]
    if len(gamma) != nph or len(pinf) != nph:
        raise ValueError("prob.EOS.gamma / pinf need %d values" % nph)
    if eos == 2:
        unsupported.append("prob.EOS.EOS = 2 (Mie-Gruneisen)")
        eos = 1
    params = make_params(dim, gamma=tuple(gamma), pinf=tuple(pinf), eos=eos,
                         b=tuple(b) if eos == 3 else None, q=tuple(q) if eos == 3 else None,
                         scheme="WENO5", riemann=riemann if riemann != "LLF" else "HLL", source_term=source_term,
                         thinc=thinc, nphase=nphase,
                         thinc_use_normal=_get(t, "FiniteVolume.THINC.UseNormal", 1, int),
                         thinc_eps=_get(t, "FiniteVolume.THINC.EPS", 1.0e-4),
                         thinc_beta=_get(t, "FiniteVolume.THINC.Beta", 2.3),
                         thinc_normal_threshold=_get(t, "FiniteVolume.THINC.NormalThreshold", 0.25))
    params.id_weno, params.id_limiter, params.id_poly, params.id_bound = w, lim, poly, bound
    # geometry and boundaries
    n_cell = _arr(t, "amr.n_cell", [], int)[:dim]
    if len(n_cell) != dim:
        raise ValueError("amr.n_cell needs %d values" % dim)
    lo = _arr(t, "geometry.prob_lo", [0.0] * 3)[:dim]
    hi = _arr(t, "geometry.prob_hi", [1.0] * 3)[:dim]
    per = _arr(t, "geometry.is_periodic", [0] * 3, int)[:dim]
    lo_bc = _arr(t, "run.lo_bc", [0] * 3, int)[:dim]
    hi_bc = _arr(t, "run.hi_bc", [0] * 3, int)[:dim]
    for d in range(dim):
        if 1 in (lo_bc[d], hi_bc[d]):
            unsupported.append("run.lo_bc / hi_bc = 1 (Inflow) in direction %d" % d)
    geom = make_geometry(tuple(n_cell), tuple(lo), tuple(hi), tuple(per), lo_bc=lo_bc, hi_bc=hi_bc)
    # time stepping, grids
    # (defaults: include/COMPAS.H:508-613; AMReX's max_grid_size default, 32 in 3D and 128 in 2D)
    ti = _get(t, "run.TimeIntegrator", "TVD-RK3", str)
    if ti not in TIME_INTEGRATORS:
        raise ValueError("run.TimeIntegrator = %r is not an integrator of the reference" % ti)
    run = dict(cfl=_get(t, "run.cfl", 0.4), timestep_change_limiter=_get(t, "run.timestep_change_limiter", 1, int),
               time_integrator=TIME_INTEGRATORS[ti],
               max_step=_get(t, "max_step", None, int), stop_time=_get(t, "stop_time", None),
               max_grid_size=_get(t, "amr.max_grid_size", 32 if dim == 3 else 128, int),
               max_level=_get(t, "amr.max_level", 0, int), ghost_cells=ng, n_cell=n_cell)
    # a hierarchy (cmp_amr_*: levels stepped together without subcycling, COMPAS.cpp:2287-2300;
    # regrid by the caller every regrid_int steps); COMPAS.H:533 defaults to subcycling
    if run["max_level"] > 0:
        ratios = _arr(t, "amr.ref_ratio", [2] * run["max_level"], int)[:run["max_level"]]
        run.update(ref_ratio=ratios, regrid_int=_get(t, "amr.regrid_int", 1, int),
                   n_error_buf=_get(t, "amr.n_error_buf", 1, int), grid_eff=_get(t, "amr.grid_eff", 0.7),
                   blocking_factor=_get(t, "amr.blocking_factor_x", _get(t, "amr.blocking_factor", 8, int), int),
                   do_subcycle=_get(t, "run.do_subcycle", 1, int))
        if run["do_subcycle"]:
            unsupported.append("run.do_subcycle = 1 (the hierarchy is stepped without subcycling, "
                               "as with run.do_subcycle = 0)")
        if len(set(ratios)) > 1 or ratios[0] not in (2, 4):
            unsupported.append("amr.ref_ratio = %s (one ratio, 2 or 4, for every level)" % ratios)
    # phase field
    pf = None
    if _get(t, "PhaseField.phase_field", 0, int):
        mech = _get(t, "PhaseField.Mechanism", "CAC-Adv", str)
        if mech not in ("CAC-Adv", "ACDI"):
            unsupported.append("PhaseField.Mechanism = %s" % mech)
            mech = "ACDI"
        erc = _get(t, "PhaseField.ID_ExplicitRC", 0, int)
        if mech == "CAC-Adv" and erc == 1:
            raise ValueError("PhaseField.Mechanism = CAC-Adv with ID_ExplicitRC = 1 (the reference aborts)")
        if _get(t, "PhaseField.ID_Flux", 0, int) != 0:
            unsupported.append("PhaseField.ID_Flux = 1 (HLL Phase-Field flux)")
        if nphase and _get(t, "PhaseField.ID_Formulation", 1, int) != 1:
            unsupported.append("PhaseField.ID_Formulation = 0 (base UD formulation)")
        for key in ("ID_DegenerateMobility", "ID_Potential", "ID_Weight"):
            if _get(t, "PhaseField." + key, 0, int):
                unsupported.append("PhaseField.%s != 0" % key)
        A, B, _, _ = eos_coefficients(eos, gamma, pinf, b if eos == 3 else None, q if eos == 3 else None)
        p_min = tuple(-B[k] / A[k] if source_term == 0 else -B[k] / (A[k] + 1.0) for k in range(nph))
        pf = make_pf_params(mechanism=mech, explicit_rc=erc,
                            flux_phase2flux=_get(t, "PhaseField.ID_FluxPhase2Flux", 1, int),
                            cfl=_get(t, "PhaseField.cfl", 0.4), lambda_=_get(t, "PhaseField.Lambda", 0.0),
                            eta=_get(t, "PhaseField.Eta", 0.0), eta_factor=_get(t, "PhaseField.Eta_Factor", 2.0),
                            eta_multiplier=_get(t, "PhaseField.Eta_Multiplier", 1.0), p_min=p_min,
                            max_step=_get(t, "PhaseField.max_step", 1000000, int),
                            change_limiter=run["timestep_change_limiter"])
    return dict(params=params, geom=geom, pf=pf, run=run, unsupported=unsupported)
