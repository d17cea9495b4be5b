"""compas_amd - MI355X (gfx950) five-equation finite-volume RHS behind a C ABI.

Python side of the library: parameter/geometry builders that map COMPAS's input keys onto the C
structs, level-buffer packing, and a Level object whose methods carry the names of the
reference's per-level operations (FillPatch, compute_dUdt_FV, ThirdOrderSSPRK, ComputeDt).
All numerics run in libcompas_amd.so; nothing here computes a flux.
"""
import ctypes

import numpy as np

from . import capi
from .capi import cmp_box, cmp_geometry, cmp_params, check

__all__ = ["Batch", "make_params", "make_pf_params", "make_geometry", "split_domain", "Layout", "FaceLayout",
           "Level", "capi", "SCHEMES", "BC"]

# FiniteVolume.Scheme -> (ID_WENO, ID_Limiter, ID_Poly, Num_Grow)
# (include/Operators_FV/FiniteVolume_Parameter.H:71-139)
SCHEMES = {
    "GODUNOV": (0, 0, 0, 1),
    "MUSCL": (0, 1, 1, 2),
    "MUSCL-Minmod": (0, 1, 1, 2),
    "MUSCL-vanLeer": (0, 2, 1, 2),
    "MUSCL-Superbee": (0, 3, 1, 2),
    "MUSCL-MC": (0, 4, 1, 2),
    "WENO5": (1, 5, 4, 3),
    "WENO5B": (1, 6, 4, 3),
}

# run.lo_bc / run.hi_bc codes (source/Class_Compressible5Eq_PhaseField.cpp:169-183)
BC = {"Interior": 0, "Inflow": 1, "Outflow": 2, "Symmetry": 3, "SlipWall": 4, "NoSlipWall": 5}

RIEMANN = {"HLL": 1, "HLLC": 2}


def eos_coefficients(eos, gamma, pinf=None, b=None, q=None):
    """A, B, C, D per phase as Parm::Initialize sets them (source/Parm.cpp:124-156, :187-272).

    eos: 0 ideal gas, 1 stiffened gas, 3 NASG (covolume b in C, heat of formation q in D)."""
    n = len(gamma)
    zero = [0.0] * n
    pinf = list(pinf) if pinf is not None else zero
    A = [1.0 / (g - 1.0) for g in gamma]
    if eos == 0:
        return A, zero, zero, zero
    Bc = [gamma[k] * pinf[k] / (gamma[k] - 1.0) for k in range(n)]
    if eos == 1:
        return A, Bc, zero, zero
    if eos == 3:
        return A, Bc, list(b) if b is not None else zero, list(q) if q is not None else zero
    raise ValueError("EOS %r not supported on this path" % (eos,))


def make_params(dim, gamma=(1.4, 1.4), pinf=None, eos=None, scheme="WENO5", riemann="HLLC",
                source_term=0, b=None, q=None, thinc=0, thinc_use_normal=1,
                thinc_eps=1.0e-4, thinc_beta=2.3, thinc_normal_threshold=0.25, nphase=0):
    """cmp_params of the two-phase five-equation model (nphase 0, two entries in gamma/pinf/b/q)
    or of the N-phase model (nphase 2 or 3, nphase entries)."""
    nph = nphase if nphase else 2
    if len(gamma) != nph:
        raise ValueError("gamma needs %d entries" % nph)
    if pinf is None:
        pinf = (0.0,) * nph
    if eos is None:
        eos = 1 if any(p != 0.0 for p in pinf) else 0
    A, B, C, D = eos_coefficients(eos, gamma, pinf, b, q)
    w, lim, poly, _ = SCHEMES[scheme]
    p = cmp_params()
    p.dim = dim
    p.nphase = nphase
    p.nstate = (2 * nphase + dim + 1) if nphase else 4 + dim
    for k in range(nph):
        p.eos_A[k], p.eos_B[k], p.eos_C[k], p.eos_D[k] = A[k], B[k], C[k], D[k]
    p.id_weno, p.id_limiter, p.id_poly, p.id_bound = w, lim, poly, 0
    p.id_thinc, p.thinc_use_normal = thinc, thinc_use_normal
    p.thinc_eps, p.thinc_beta, p.thinc_normal_threshold = thinc_eps, thinc_beta, thinc_normal_threshold
    p.source_term = source_term
    p.id_riemann = RIEMANN[riemann]
    return p


def scheme_ghosts(scheme):
    return SCHEMES[scheme][3]


# PhaseField.Mechanism strings with a path here -> ID_Mechanism
# (include/Physics_PhaseField/PhaseField_Parameter.H:79-137)
PF_MECHANISMS = {"CAC-Adv": 1, "ACDI": 3}


def make_pf_params(mechanism="ACDI", explicit_rc=1, id_flux=0, flux_phase2flux=1, cfl=0.4,
                   lambda_=0.0, eta=0.0, eta_factor=2.0, eta_multiplier=1.0, p_min=(0.0, 0.0),
                   max_step=1000000, change_limiter=1):
    """cmp_pf_params with PhaseField_Parameter's defaults (PhaseField_Parameter.H:31-53)."""
    p = capi.cmp_pf_params()
    p.mechanism = PF_MECHANISMS[mechanism]
    p.explicit_rc, p.id_flux, p.flux_phase2flux = explicit_rc, id_flux, flux_phase2flux
    p.cfl, p.lambda_, p.eta, p.eta_factor, p.eta_multiplier = cfl, lambda_, eta, eta_factor, eta_multiplier
    if len(p_min) > 3:
        raise ValueError("p_min holds at most 3 phases")
    for k, v in enumerate(p_min):
        p.p_min[k] = v
    p.max_step, p.change_limiter = max_step, change_limiter
    return p


def make_geometry(n_cell, prob_lo, prob_hi, periodic, lo_bc=None, hi_bc=None):
    dim = len(n_cell)
    g = cmp_geometry()
    g.dim = dim
    for d in range(3):
        if d < dim:
            g.domain_lo[d] = 0
            g.domain_hi[d] = n_cell[d] - 1
            g.is_periodic[d] = int(periodic[d])
            g.dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d]
            g.prob_lo[d] = prob_lo[d]
            g.lo_bc[d] = 0 if periodic[d] else (lo_bc[d] if lo_bc else BC["Outflow"])
            g.hi_bc[d] = 0 if periodic[d] else (hi_bc[d] if hi_bc else BC["Outflow"])
        else:
            g.domain_lo[d] = g.domain_hi[d] = 0
            g.dx[d] = 1.0
    return g


def split_domain(n_cell, max_grid_size):
    """Uniform split of the domain into boxes of at most max_grid_size cells a side
    (AMReX's maxSize chopping of a single-level BoxArray)."""
    dim = len(n_cell)
    mg = max_grid_size if isinstance(max_grid_size, (list, tuple)) else [max_grid_size] * dim
    edges = []
    for d in range(dim):
        nb = -(-n_cell[d] // mg[d])
        base, extra = divmod(n_cell[d], nb)
        cuts = [0]
        for i in range(nb):
            cuts.append(cuts[-1] + base + (1 if i < extra else 0))
        edges.append(cuts)
    boxes = []
    ranges = [range(len(e) - 1) for e in edges]
    import itertools
    for idx in itertools.product(*reversed(ranges)):
        idx = tuple(reversed(idx))
        lo = [edges[d][idx[d]] for d in range(dim)]
        hi = [edges[d][idx[d] + 1] - 1 for d in range(dim)]
        boxes.append((lo, hi))
    return boxes


class Layout:
    """Level-buffer layout: boxes with ng ghosts, ncomp components, fabs concatenated."""

    def __init__(self, dim, boxes, ng, ncomp):
        self.dim, self.ng, self.ncomp = dim, ng, ncomp
        self.nbox = len(boxes)
        self.cboxes = (cmp_box * self.nbox)()
        for b, (lo, hi) in enumerate(boxes):
            for d in range(3):
                self.cboxes[b].lo[d] = lo[d] if d < dim else 0
                self.cboxes[b].hi[d] = hi[d] if d < dim else 0
        total = ctypes.c_longlong(0)
        lib = capi.load()
        check(lib.cmp_level_layout(dim, self.nbox, self.cboxes, ng, ncomp, ctypes.byref(total)),
              "cmp_level_layout")
        self.total = total.value
        self.boxes = [(list(lo), list(hi)) for lo, hi in boxes]

    def fab_shape(self, b):
        lo, hi = self.boxes[b]
        ext = [hi[d] - lo[d] + 1 + 2 * self.ng for d in range(self.dim)]
        return (self.ncomp,) + tuple(reversed(ext))

    def fab_view(self, buf, b):
        """numpy view of box b's fab in a level buffer: [n, (k,) j, i]."""
        off = self.cboxes[b].offset
        shape = self.fab_shape(b)
        size = int(np.prod(shape))
        return buf[off:off + size].reshape(shape)

    def pack(self, field, buf=None):
        """Global valid field [n, (k,) j, i] -> level buffer (ghosts zero)."""
        if buf is None:
            buf = np.zeros(self.total, dtype=np.float64)
        g = self.ng
        for b, (lo, hi) in enumerate(self.boxes):
            v = self.fab_view(buf, b)
            sl_dst = (slice(None),) + tuple(slice(g, g + hi[d] - lo[d] + 1) for d in reversed(range(self.dim)))
            sl_src = (slice(None),) + tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(self.dim)))
            v[sl_dst] = field[sl_src]
        return buf

    def unpack(self, buf, shape):
        """Level buffer -> global valid field [n, (k,) j, i]."""
        out = np.empty((self.ncomp,) + tuple(shape), dtype=np.float64)
        g = self.ng
        for b, (lo, hi) in enumerate(self.boxes):
            v = self.fab_view(buf, b)
            sl_src = (slice(None),) + tuple(slice(g, g + hi[d] - lo[d] + 1) for d in reversed(range(self.dim)))
            sl_dst = (slice(None),) + tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(self.dim)))
            out[sl_dst] = v[sl_src]
        return out


class FaceLayout:
    """Face-buffer layout of a level (cmp_face_layout): per box and direction a nodal array of
    ncomp components, box-major."""

    def __init__(self, dim, boxes, ncomp=1):
        self.dim, self.ncomp = dim, ncomp
        self.boxes = [(list(lo), list(hi)) for lo, hi in boxes]
        self.offsets = []
        off = 0
        for lo, hi in self.boxes:
            row = []
            for d in range(dim):
                row.append(off)
                off += ncomp * int(np.prod([hi[e] - lo[e] + 1 + (e == d) for e in range(dim)]))
            self.offsets.append(row)
        self.total = off

    def face_shape(self, b, d):
        lo, hi = self.boxes[b]
        ext = [hi[e] - lo[e] + 1 + (e == d) for e in range(self.dim)]
        return (self.ncomp,) + tuple(reversed(ext))

    def face_view(self, buf, b, d):
        """numpy view of box b's direction-d faces: [n, (k,) j, i]."""
        shape = self.face_shape(b, d)
        off = self.offsets[b][d]
        return buf[off:off + int(np.prod(shape))].reshape(shape)


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(stream.cuda_stream)


def tile_cache_save(path):
    """The tile autotuner's cache as a text file."""
    check(capi.load().cmp_tile_cache_save(str(path).encode()), "cmp_tile_cache_save")


def tile_cache_load(path):
    check(capi.load().cmp_tile_cache_load(str(path).encode()), "cmp_tile_cache_load")


def tile_cache_clear():
    check(capi.load().cmp_tile_cache_clear(), "cmp_tile_cache_clear")


class Level:
    """One level's boxes on the GPU. Buffers are torch CUDA tensors owned by the caller."""

    def __init__(self, params, geom, layout):
        self.lib = capi.load()
        self.params, self.geom, self.layout = params, geom, layout
        h = ctypes.c_void_p()
        check(self.lib.cmp_level_create(ctypes.byref(params), ctypes.byref(geom), layout.nbox,
                                        layout.cboxes, layout.ng, ctypes.byref(h)), "cmp_level_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.cmp_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # FillPatchSingleLevel
    def FillPatch(self, U, stream=None):
        check(self.lib.cmp_level_fill_ghosts(self.h, _ptr(U), _stream(stream)), "cmp_level_fill_ghosts")

    # compute_dUdt_FV on every box (ghosts of U filled)
    def compute_dUdt_FV(self, dUdt, U, cmax=None, stream=None):
        check(self.lib.cmp_level_rhs(self.h, _ptr(U), _ptr(dUdt), _ptr(cmax), _stream(stream)),
              "cmp_level_rhs")

    def rk3_stage(self, stage, Us, Un, Uout, dt_dev, cmax=None, stream=None):
        check(self.lib.cmp_level_rk3_stage(self.h, stage, _ptr(Us), _ptr(Un), _ptr(Uout), _ptr(dt_dev),
                                           _ptr(cmax), _stream(stream)), "cmp_level_rk3_stage")

    def ThirdOrderSSPRK(self, Un, U1, U2, dt_dev, cmax=None, stream=None):
        check(self.lib.cmp_level_rk3_step(self.h, _ptr(Un), _ptr(U1), _ptr(U2), _ptr(dt_dev),
                                          _ptr(cmax), _stream(stream)), "cmp_level_rk3_step")

    def CalculateInitialcmax(self, U, cmax, stream=None):
        check(self.lib.cmp_level_initial_cmax(self.h, _ptr(U), _ptr(cmax), _stream(stream)),
              "cmp_level_initial_cmax")

    def ComputeDt(self, cmax, dt_dev, cfl, limiter=1, stream=None):
        check(self.lib.cmp_level_compute_dt(self.h, _ptr(cmax), _ptr(dt_dev), float(cfl), int(limiter),
                                            _stream(stream)), "cmp_level_compute_dt")

    def advance(self, Un, U1, U2, nsteps, cfl, dt_dev, cmax, dt_hist=None, limiter=1, use_graph=True,
                resume=False, stream=None):
        check(self.lib.cmp_level_advance(self.h, _ptr(Un), _ptr(U1), _ptr(U2), int(nsteps), float(cfl),
                                         int(limiter), _ptr(dt_dev), _ptr(cmax), _ptr(dt_hist),
                                         int(use_graph), int(bool(resume)), _stream(stream)),
              "cmp_level_advance")

    # the level's clock (advanced by every ComputeDt) and stop_time (negative: none)
    def set_time(self, time, stop_time=-1.0, stream=None):
        check(self.lib.cmp_level_set_time(self.h, float(time), float(stop_time), _stream(stream)),
              "cmp_level_set_time")

    def get_time(self, stream=None):
        t = ctypes.c_double()
        check(self.lib.cmp_level_get_time(self.h, ctypes.byref(t), _stream(stream)), "cmp_level_get_time")
        return t.value

    # run.TimeIntegrator of advance() and the Phase-Field sub-steps (1..4)
    def set_integrator(self, integrator):
        check(self.lib.cmp_level_set_integrator(self.h, int(integrator)), "cmp_level_set_integrator")

    # one step of run.TimeIntegrator: 1 ForwardEuler, 2 TVD-RK2, 3 TVD-RK3, 4 RK4 (U3 for RK4 only)
    def rk_step(self, integrator, Un, U1, U2, dt_dev, cmax, U3=None, stream=None):
        check(self.lib.cmp_level_rk_step(self.h, int(integrator), _ptr(Un), _ptr(U1), _ptr(U2), _ptr(U3),
                                         _ptr(dt_dev), _ptr(cmax), _stream(stream)), "cmp_level_rk_step")

    # one Phase-Field sub-step of pf_dt at a fixed Grad_Q with run.TimeIntegrator
    def pf_rk_step(self, integrator, Un, U1, U2, gradq, pf_dt, U3=None, stream=None):
        check(self.lib.cmp_level_pf_rk_step(self.h, int(integrator), _ptr(Un), _ptr(U1), _ptr(U2), _ptr(U3),
                                            _ptr(gradq), float(pf_dt), _stream(stream)), "cmp_level_pf_rk_step")

    # AverageDownTo: this (fine) level's Uf onto the coarse level's Uc under the fine boxes
    def AverageDownTo(self, Uf, coarse, Uc, ratio, stream=None):
        check(self.lib.cmp_level_average_down(self.h, _ptr(Uf), coarse.h, _ptr(Uc), int(ratio), _stream(stream)),
              "cmp_level_average_down")

    # FillPatchTwoLevels: this (fine) level's ghost cells from its own boxes and, where none holds
    # the source, the coarse level's conservative linear interpolation
    def FillPatchTwoLevels(self, Uf, coarse, Uc, ratio, stream=None):
        check(self.lib.cmp_level_fill_two_levels(self.h, _ptr(Uf), coarse.h, _ptr(Uc), int(ratio),
                                                 _stream(stream)), "cmp_level_fill_two_levels")

    # RemakeLevel / MakeNewLevelFromCoarse: every cell of this (new) level from the old level's boxes
    # (old None: none) where they hold it, else the coarse level's interpolation
    def FillFrom(self, U, old, Uold, coarse, Uc, ratio, stream=None):
        check(self.lib.cmp_level_fill_from(self.h, _ptr(U), old.h if old is not None else None, _ptr(Uold),
                                           coarse.h, _ptr(Uc), int(ratio), _stream(stream)), "cmp_level_fill_from")

    # ErrorEst: tags (torch uint8, one byte per valid cell, box by box) from a ghost-filled U
    def ErrorEst(self, U, criteria, tags, stream=None):
        arr = tag_criteria(criteria)
        check(self.lib.cmp_level_tag(self.h, _ptr(U), arr, len(criteria), _ptr(tags), _stream(stream)),
              "cmp_level_tag")

    def check_errors(self, stream=None):
        """Synchronises and returns the device error word (0 when clean)."""
        return self.lib.cmp_level_check(self.h, _stream(stream))

    def info(self):
        n, x, y, z = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int())
        check(self.lib.cmp_level_info(self.h, ctypes.byref(n), ctypes.byref(x), ctypes.byref(y),
                                      ctypes.byref(z)), "cmp_level_info")
        return {"ntiles": n.value, "tile": (x.value, y.value, z.value)}

    # ---- the persistent work-list stage kernel and the tile autotuner ------------------------------
    def set_tiles(self, ext, waves=2):
        """The TVD-RK3 flow stage through the work-list kernel, box b in tiles of ext[b] (3 ints per
        box; one triple for every box; 0 0 0: the box stays with the fixed-shape kernel), occupancy
        target `waves`; ext None: the fixed-shape kernel."""
        if ext is None:
            check(self.lib.cmp_level_set_tiles(self.h, None, 0), "cmp_level_set_tiles")
            return
        e = [int(v) for v in ext] if len(ext) != 3 else [int(v) for v in ext] * self.layout.nbox
        arr = (ctypes.c_int * (3 * self.layout.nbox))(*e)
        check(self.lib.cmp_level_set_tiles(self.h, arr, int(waves)), "cmp_level_set_tiles")

    def tile_shapes(self):
        """(work-list kernel on, its occupancy target, [tile extents of each box]; (0, 0, 0) for a
        box the fixed-shape kernel keeps in a level that mixes the two)"""
        on, w = ctypes.c_int(), ctypes.c_int()
        arr = (ctypes.c_int * (3 * self.layout.nbox))()
        check(self.lib.cmp_level_tile_shapes(self.h, ctypes.byref(on), ctypes.byref(w), arr),
              "cmp_level_tile_shapes")
        return bool(on.value), w.value, [tuple(arr[3 * b:3 * b + 3]) for b in range(self.layout.nbox)]

    def tune_tiles(self, U, extra_shapes=(), force=0, stream=None):
        """The measured tile-shape sweep for this level's box shapes (and extra_shapes), then the
        level's choice (force 1 / -1: work-list kernel always / never)."""
        ex = [int(v) for sh in extra_shapes for v in (list(sh) + [1, 1])[:3]]
        arr = (ctypes.c_int * max(1, len(ex)))(*ex)
        check(self.lib.cmp_level_tune_tiles(self.h, _ptr(U), len(ex) // 3, arr if ex else None, int(force),
                                            _stream(stream)), "cmp_level_tune_tiles")

    def tile_tuning(self, shape):
        """The cached sweep of one box shape: {"fixed": ns per cell, "cand": [(ext, ns at 2 waves,
        ns at 3 waves)]}, or None when not swept."""
        n = (ctypes.c_int * 3)(*(list(shape) + [1, 1])[:3])
        nf, nc = ctypes.c_double(), ctypes.c_int()
        ce = (ctypes.c_int * 36)()
        cn = (ctypes.c_double * 24)()
        if self.lib.cmp_tile_tuning(self.h, n, ctypes.byref(nf), ctypes.byref(nc), ce, cn) != 0:
            return None
        return {"fixed": nf.value,
                "cand": [(tuple(ce[3 * k:3 * k + 3]), cn[2 * k], cn[2 * k + 1]) for k in range(nc.value)]}

    # ---- Phase-Field regularization ----------------------------------------------------------
    def pf_setup(self, pf):
        self.pf = pf
        # Grad_Q: one component per volume fraction
        self.faces = FaceLayout(self.layout.dim, self.layout.boxes, self.params.nphase or 1)
        check(self.lib.cmp_level_pf_setup(self.h, ctypes.byref(pf)), "cmp_level_pf_setup")

    # PhaseField_compute_dUdt_FV on every box (ghosts of U filled; gradq a face buffer)
    def PhaseField_compute_dUdt_FV(self, dUdt, U, gradq, stream=None):
        check(self.lib.cmp_level_pf_rhs(self.h, _ptr(U), _ptr(gradq), _ptr(dUdt), _stream(stream)),
              "cmp_level_pf_rhs")

    # MaxVelocity + PhaseFieldMechanism_Evolve + PhaseField_GradQPF; returns (max|Grad_Q|, info)
    def PhaseField_GradQPF(self, U, dt, gradq, stream=None):
        gm = ctypes.c_double(0.0)
        info = (ctypes.c_double * 4)()
        check(self.lib.cmp_level_pf_gradq(self.h, _ptr(U), float(dt), _ptr(gradq), ctypes.byref(gm),
                                          info, _stream(stream)), "cmp_level_pf_gradq")
        return gm.value, list(info)

    def solver_iterations(self):
        """Iterations of the last implicit Grad_Q solve."""
        it = ctypes.c_int(0)
        check(self.lib.cmp_level_pf_solver_iterations(self.h, ctypes.byref(it)),
              "cmp_level_pf_solver_iterations")
        return it.value

    def PhaseField_ThirdOrderSSPRK(self, Un, U1, U2, gradq, pf_dt, stream=None):
        check(self.lib.cmp_level_pf_rk3_step(self.h, _ptr(Un), _ptr(U1), _ptr(U2), _ptr(gradq),
                                             float(pf_dt), _stream(stream)), "cmp_level_pf_rk3_step")

    def PhaseField_Evolve(self, Un, U1, U2, dt_dev, stream=None):
        """Returns [max|u|, mobility, eta, mechanism sub-steps, max|Grad_Q|, Phase-Field sub-steps]."""
        info = (ctypes.c_double * 6)()
        check(self.lib.cmp_level_pf_evolve(self.h, _ptr(Un), _ptr(U1), _ptr(U2), _ptr(dt_dev), info,
                                           _stream(stream)), "cmp_level_pf_evolve")
        return list(info)


TAG_VARS = {"rho": capi.CMP_TAG_RHO, "P": capi.CMP_TAG_P, "u": capi.CMP_TAG_U, "v": capi.CMP_TAG_V,
            "w": capi.CMP_TAG_W}
TAG_KINDS = {"grad": capi.CMP_TAG_GRAD, "value_less": capi.CMP_TAG_LESS, "value_greater": capi.CMP_TAG_GREATER,
             "value_range": capi.CMP_TAG_RANGE}


def tag_criteria(criteria):
    """[(var, kind, v0[, v1])] -> a ctypes array of cmp_tag_criterion; var is a state component
    index or one of the derived names of TAG_VARS, kind one of TAG_KINDS (the run.refine.<var>.*
    keys)."""
    arr = (capi.cmp_tag_criterion * max(1, len(criteria)))()
    for q, c in enumerate(criteria):
        var, kind, v0 = c[0], c[1], c[2]
        arr[q].var = TAG_VARS[var] if isinstance(var, str) else int(var)
        arr[q].kind = TAG_KINDS[kind]
        arr[q].v0 = float(v0)
        arr[q].v1 = float(c[3]) if len(c) > 3 else 0.0
    return arr


def rhs_box(params, geom, box, ng, U, dUdt, cmax=None, fluxes=(None, None, None),
            fluxnc=(None, None, None), stream=None):
    """Per-box compute_dUdt_FV through cmp_rhs_box (torch CUDA tensors)."""
    lib = capi.load()
    b = cmp_box()
    lo, hi = box
    for d in range(3):
        b.lo[d] = lo[d] if d < len(lo) else 0
        b.hi[d] = hi[d] if d < len(hi) else 0
    check(lib.cmp_rhs_box(ctypes.byref(params), ctypes.byref(geom), ctypes.byref(b), ng, _ptr(U),
                          _ptr(dUdt), _ptr(cmax), _ptr(fluxes[0]), _ptr(fluxes[1]), _ptr(fluxes[2]),
                          _ptr(fluxnc[0]), _ptr(fluxnc[1]), _ptr(fluxnc[2]), _stream(stream)),
          "cmp_rhs_box")


def cmax_box(params, geom, box, ng, U, cmax_cells, stream=None):
    """Per-cell c_max of one box through cmp_cmax_box (torch CUDA tensors)."""
    lib = capi.load()
    b = cmp_box()
    lo, hi = box
    for d in range(3):
        b.lo[d] = lo[d] if d < len(lo) else 0
        b.hi[d] = hi[d] if d < len(hi) else 0
    check(lib.cmp_cmax_box(ctypes.byref(params), ctypes.byref(geom), ctypes.byref(b), ng, _ptr(U), _ptr(cmax_cells),
                           _stream(stream)), "cmp_cmax_box")


def pf_rhs_box(params, pf, geom, box, ng, U, gradq, dUdt, stream=None):
    """Per-box PhaseField_compute_dUdt_FV through cmp_pf_rhs_box (torch CUDA tensors; gradq in the
    box's own cmp_face_layout, one component per volume fraction)."""
    lib = capi.load()
    b = cmp_box()
    lo, hi = box
    for d in range(3):
        b.lo[d] = lo[d] if d < len(lo) else 0
        b.hi[d] = hi[d] if d < len(hi) else 0
    check(lib.cmp_pf_rhs_box(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom), ctypes.byref(b), ng,
                             _ptr(U), _ptr(gradq), _ptr(dUdt), _stream(stream)), "cmp_pf_rhs_box")


class Amr:
    """A hierarchy of Levels (coarsest first, each the refinement of the previous by ratio) stepped
    together without subcycling (cmp_amr_*): every RK stage fills every level from the stage state,
    gives the coarse faces under a finer level the finer level's face means, and averages down.
    Level buffers (torch CUDA tensors, one per level) stay the caller's."""

    def __init__(self, levels, ratio=2):
        self.lib = capi.load()
        self.levels = list(levels)
        self.ratio = int(ratio)
        n = len(self.levels)
        arr = (ctypes.c_void_p * n)(*[lv.h.value for lv in self.levels])
        h = ctypes.c_void_p()
        check(self.lib.cmp_amr_create(arr, n, self.ratio, ctypes.byref(h)), "cmp_amr_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.cmp_amr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _arr(ts):
        return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    def tiles(self, force=0):
        """Every level's flow stage in one work-list launch (force 1 on, -1 off, 0 by the tuned times);
        returns (on, occupancy target)."""
        on, w = ctypes.c_int(), ctypes.c_int()
        check(self.lib.cmp_amr_tiles(self.h, int(force), ctypes.byref(on), ctypes.byref(w)), "cmp_amr_tiles")
        return bool(on.value), w.value

    def tile_info(self):
        """(one work-list launch for every level, its occupancy target), changing nothing."""
        on, w = ctypes.c_int(), ctypes.c_int()
        check(self.lib.cmp_amr_tile_info(self.h, ctypes.byref(on), ctypes.byref(w)), "cmp_amr_tile_info")
        return bool(on.value), w.value

    def cf_cells(self, level):
        n = ctypes.c_int()
        check(self.lib.cmp_amr_info(self.h, int(level), ctypes.byref(n)), "cmp_amr_info")
        return n.value

    def initial_cmax(self, U, cmax, stream=None):
        check(self.lib.cmp_amr_initial_cmax(self.h, self._arr(U), _ptr(cmax), _stream(stream)),
              "cmp_amr_initial_cmax")

    def compute_dt(self, cmax, dt_dev, cfl, limiter=1, stream=None):
        check(self.lib.cmp_amr_compute_dt(self.h, _ptr(cmax), _ptr(dt_dev), float(cfl), int(limiter),
                                          _stream(stream)), "cmp_amr_compute_dt")

    def rk3_step(self, Un, U1, U2, dt_dev, cmax, stream=None):
        check(self.lib.cmp_amr_rk3_step(self.h, self._arr(Un), self._arr(U1), self._arr(U2), _ptr(dt_dev),
                                        _ptr(cmax), _stream(stream)), "cmp_amr_rk3_step")

    def advance(self, Un, U1, U2, nsteps, cfl, dt_dev, cmax, dt_hist=None, limiter=1, use_graph=True,
                stream=None):
        check(self.lib.cmp_amr_advance(self.h, self._arr(Un), self._arr(U1), self._arr(U2), int(nsteps),
                                       float(cfl), int(limiter), _ptr(dt_dev), _ptr(cmax), _ptr(dt_hist),
                                       int(use_graph), _stream(stream)), "cmp_amr_advance")

    # the Phase-Field step of the hierarchy (PhaseField_Evolve over the levels)
    def pf_setup(self, pf):
        check(self.lib.cmp_amr_pf_setup(self.h, ctypes.byref(pf)), "cmp_amr_pf_setup")

    def PhaseField_GradQPF(self, U, dt, gradq=None, stream=None):
        """Grad_Q of every level (into the torch face buffers gradq[l] when given); returns
        (max|Grad_Q| per level, [max|u|, mobility, eta, mechanism sub-steps, CG iterations])."""
        n = len(self.levels)
        gm = (ctypes.c_double * n)()
        info = (ctypes.c_double * 5)()
        g = self._arr(gradq) if gradq is not None else None
        check(self.lib.cmp_amr_pf_gradq(self.h, self._arr(U), float(dt), g, gm, info, _stream(stream)),
              "cmp_amr_pf_gradq")
        return list(gm), list(info)

    def PhaseField_Evolve(self, Un, U1, U2, dt_dev, stream=None):
        """[max|u|, mobility, eta, mechanism sub-steps, max|Grad_Q|, Phase-Field sub-steps, CG iterations]"""
        info = (ctypes.c_double * 7)()
        check(self.lib.cmp_amr_pf_evolve(self.h, self._arr(Un), self._arr(U1), self._arr(U2), _ptr(dt_dev), info,
                                         _stream(stream)), "cmp_amr_pf_evolve")
        return list(info)

    def solver_iterations(self):
        it = ctypes.c_int()
        check(self.lib.cmp_amr_pf_solver_iterations(self.h, ctypes.byref(it)), "cmp_amr_pf_solver_iterations")
        return it.value

    def check_errors(self, stream=None):
        return max(lv.check_errors(stream=stream) for lv in self.levels)


class Batch:
    """Independent levels ("cases") advanced together (cmp_batch_*): each admitted case keeps its
    state in device memory; step() advances every case by one RK3 step with its own dt, in one
    launch per stage over the tiles of all cases."""

    def __init__(self, like, capacity_doubles, max_cases):
        self.lib = capi.load()
        h = ctypes.c_void_p()
        check(self.lib.cmp_batch_create(ctypes.byref(like), int(capacity_doubles), int(max_cases),
                                        ctypes.byref(h)), "cmp_batch_create")
        self.h = h
        self.max_cases = int(max_cases)
        self.layouts = {}

    def close(self):
        if self.h:
            self.lib.cmp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, params, geom, layout, U0, cfl, limiter=1, stream=None):
        """Admits a case (U0: host level buffer of `layout`); returns its id, or None when the batch
        has no free slot or arena range for it (CMP_ERR_NOMEM)."""
        buf = np.ascontiguousarray(U0, dtype=np.float64)
        cid = ctypes.c_int(-1)
        rc = self.lib.cmp_batch_admit(self.h, ctypes.byref(params), ctypes.byref(geom), layout.nbox,
                                      layout.cboxes, layout.ng,
                                      buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(cfl),
                                      int(limiter), ctypes.byref(cid), _stream(stream))
        if rc == capi.CMP_ERR_NOMEM:
            return None
        check(rc, "cmp_batch_admit")
        self.layouts[cid.value] = layout
        return cid.value

    def retire(self, case_id, stream=None):
        check(self.lib.cmp_batch_retire(self.h, int(case_id), _stream(stream)), "cmp_batch_retire")
        self.layouts.pop(case_id, None)

    def step(self, stream=None):
        check(self.lib.cmp_batch_step(self.h, _stream(stream)), "cmp_batch_step")

    def pf_setup(self, pf):
        """The Phase-Field step for every admitted and every later case (cmp_batch_pf_setup)."""
        check(self.lib.cmp_batch_pf_setup(self.h, ctypes.byref(pf)), "cmp_batch_pf_setup")

    def pf_step(self, nthreads=0, stream=None):
        """The Phase-Field step of every admitted case after step(); returns {case id: info} with
        cmp_level_pf_evolve's six numbers per case."""
        info = (ctypes.c_double * (6 * self.max_cases))()
        check(self.lib.cmp_batch_pf_step(self.h, int(nthreads), info, _stream(stream)), "cmp_batch_pf_step")
        return info

    def state(self, case_id, stream=None):
        """(host copy of the case's U^n level buffer, simulated time, steps taken)."""
        ptr, n = ctypes.c_void_p(), ctypes.c_longlong(0)
        t, k = ctypes.c_double(0.0), ctypes.c_int(0)
        check(self.lib.cmp_batch_case_state(self.h, int(case_id), ctypes.byref(ptr), ctypes.byref(n),
                                            ctypes.byref(t), ctypes.byref(k), _stream(stream)),
              "cmp_batch_case_state")
        out = np.empty(n.value, dtype=np.float64)
        check(self.lib.cmp_batch_case_download(self.h, int(case_id),
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               _stream(stream)), "cmp_batch_case_download")
        return out, t.value, k.value

    def check_errors(self, case_id, stream=None):
        return self.lib.cmp_batch_check(self.h, int(case_id), _stream(stream))

    def info(self):
        n, t, f = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_longlong(0)
        check(self.lib.cmp_batch_info(self.h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(f)), "cmp_batch_info")
        return n.value, t.value, f.value
