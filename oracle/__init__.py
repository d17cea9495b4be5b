"""The oracle: ctypes binding of oracle/_build/libcpu_ref.so (CPU restatement of COMPAS's
five-equation hyperbolic RHS).

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
may import this package, and only as the checker (or as the timed CPU baseline); the product
path never does. numpy arrays in, numpy arrays out.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libcpu_ref.so")

_lib = None
_lib_path = None     # another build of the same source (the bench's -march=native baseline), or None


def _dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def use_library(path=None):
    """Binds another build of oracle/cpu_ref.c (path), or the default one (None), from the next call
    on. The bench's CPU baseline times a -march=native build this way."""
    global _lib, _lib_path
    _lib, _lib_path = None, path


def load():
    global _lib
    if _lib is not None:
        return _lib
    if _lib_path is None and not os.path.exists(LIB_PATH):
        from . import build as _b
        _b.build()
    # the product library's struct mirrors are shared (plain PODs)
    from compas_amd.capi import cmp_box, cmp_geometry, cmp_params, cmp_pf_params, cmp_tag_criterion
    lib = ctypes.CDLL(_lib_path or LIB_PATH)
    D = ctypes.c_double
    PD = ctypes.POINTER(D)
    PI = ctypes.POINTER(ctypes.c_int)
    PP = ctypes.POINTER(cmp_params)
    PF = ctypes.POINTER(cmp_pf_params)
    PG = ctypes.POINTER(cmp_geometry)
    PB = ctypes.POINTER(cmp_box)
    sigs = {
        "ref_nasg_phi": (D, [D, D, D, PI]),
        "ref_eos_mixture": (D, [PP, D, D, D, D, ctypes.c_int, PI]),
        "ref_cons2prim": (None, [PP, PD, PD, PI]),
        "ref_prim2cons": (None, [PP, PD, PD, PI]),
        "ref_rho_c2_phase": (D, [D, D, D]),
        "ref_sound_speed": (D, [PP, D, D, D, D, ctypes.c_int, PI]),
        "ref_nonconservative_K": (None, [PP, D, D, D, D, ctypes.c_int, PD, PI]),
        "ref_cons2flux": (None, [PP, PD, ctypes.c_int, PD, PI]),
        "ref_weno5": (D, [D, D, D, D, D, ctypes.c_int]),
        "ref_weno5_weights": (None, [D, D, D, D, D, ctypes.c_int, PD, PD]),
        "ref_muscl": (D, [D, D, D, D, ctypes.c_int, ctypes.c_int]),
        "ref_weno5b_side": (None, [PP, PD, ctypes.c_int, PD]),
        "ref_thinc": (D, [D, D, D, ctypes.c_int]),
        "ref_riemann": (None, [PP, PD, PD, ctypes.c_int, PD, PD, PD, PI]),
        "ref_np_eos_mixture": (D, [PP, D, PD, PD, ctypes.c_int, PI]),
        "ref_np_cons2prim": (None, [PP, PD, PD, PI]),
        "ref_np_prim2cons": (None, [PP, PD, PD, PI]),
        "ref_np_sound_speed": (D, [PP, D, PD, PD, ctypes.c_int, PI]),
        "ref_np_K": (None, [PP, D, PD, PD, ctypes.c_int, PD, PI]),
        "ref_np_cons2flux": (None, [PP, PD, ctypes.c_int, PD, PI]),
        "ref_np_riemann": (None, [PP, PD, PD, ctypes.c_int, PD, PD, PD, PI]),
        "ref_np_limiter_consistent": (None, [ctypes.c_int, PD, PD]),
        "ref_np_pf_flux": (None, [PP, PF, PD, PD, PD, PD, PD, PI]),
        "ref_np_cons2phase": (None, [PP, PF, PD, PD, PI]),
        "ref_np_limit_consistent": (None, [PP, PD, PD, D, ctypes.c_int, PD, PI]),
        "ref_face_flux": (ctypes.c_int, [PP, PG, PB, ctypes.c_int, PD, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, PD, PD, PD]),
        "ref_rhs_box": (ctypes.c_int, [PP, PG, PB, ctypes.c_int, PD, PD, PD,
                                       ctypes.POINTER(PD), ctypes.POINTER(PD)]),
        "ref_level_fill_ghosts": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD]),
        "ref_level_rhs": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD, PD, PD]),
        "ref_level_rk3_step": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD, D, PD]),
        "ref_level_initial_cmax": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD, PD]),
        "ref_est_dt": (D, [PG, D, D]),
        "ref_compute_dt": (D, [PG, D, D, ctypes.c_int, D]),
        "ref_level_advance": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD,
                                             ctypes.c_int, D, ctypes.c_int, PD]),
        "ref_num_threads": (ctypes.c_int, []),
        "ref_set_num_threads": (None, [ctypes.c_int]),
        "ref_level_fill_pf": (ctypes.c_int, [ctypes.c_int, PG, ctypes.c_int, PB, ctypes.c_int,
                                             ctypes.c_int, PD]),
        "ref_pf_map_backward": (D, [D, D]),
        "ref_pf_map_forward": (D, [D, D, ctypes.c_int]),
        "ref_pf_acdi_flux": (D, [D, D, D, D, D, D, D]),
        "ref_pf_gradq_face": (D, [D, D, D, D, D]),
        "ref_pf_flux": (None, [PP, PF, PD, PD, D, PD, PD, PI]),
        "ref_pf_coefficients": (None, [PF, PG, D, PD, PD, PD]),
        "ref_pf_compute_dt": (D, [PF, PG, D, D, D, PD]),
        "ref_level_pf_rhs": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, ctypes.c_int, PD, PD, PD]),
        "ref_level_pf_rk3_step": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, ctypes.c_int, PD, PD, D]),
        "ref_level_max_velocity": (D, [PP, ctypes.c_int, PB, ctypes.c_int, PD]),
        "ref_level_average_down": (ctypes.c_int, [PP, ctypes.c_int, PB, ctypes.c_int, PD, ctypes.c_int, PB,
                                                  ctypes.c_int, PD, ctypes.c_int]),
        "ref_level_fill_two_levels": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD, PG, ctypes.c_int,
                                                     PB, ctypes.c_int, PD, ctypes.c_int]),
        "ref_level_fill_from": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD, ctypes.c_int, PB,
                                               ctypes.c_int, PD, PG, ctypes.c_int, PB, ctypes.c_int, PD, ctypes.c_int]),
        "ref_level_rk_step": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD, D, PD, ctypes.c_int]),
        "ref_level_advance_int": (ctypes.c_int, [PP, PG, ctypes.c_int, PB, ctypes.c_int, PD, ctypes.c_int, D,
                                                 ctypes.c_int, PD, ctypes.c_int]),
        "ref_level_advance_pf_int": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, ctypes.c_int, PD, ctypes.c_int, D,
                                                    ctypes.c_int, PD, ctypes.c_int]),
        "ref_level_pf_rk_step": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, ctypes.c_int, PD, PD, D, ctypes.c_int]),
        "ref_amr_rk3_step": (ctypes.c_int, [PP, ctypes.c_int, PG, PI, ctypes.POINTER(PB), ctypes.c_int,
                                            ctypes.POINTER(PD), D, ctypes.c_int, PD]),
        "ref_amr_compute_dt": (D, [ctypes.c_int, PG, PD, D, ctypes.c_int, D]),
        "ref_amr_pf_gradq_step": (ctypes.c_int, [PP, PF, ctypes.c_int, PG, PI, ctypes.POINTER(PB), ctypes.c_int,
                                                 ctypes.POINTER(PD), D, ctypes.c_int, ctypes.POINTER(PD), PD]),
        "ref_amr_pf_gradq_implicit": (ctypes.c_int, [PP, PF, ctypes.c_int, PG, PI, ctypes.POINTER(PB),
                                                     ctypes.POINTER(PD), ctypes.POINTER(PD), D, ctypes.c_int,
                                                     ctypes.POINTER(PD), ctypes.POINTER(PD), PI]),
        "ref_amr_pf_evolve": (ctypes.c_int, [PP, PF, ctypes.c_int, PG, PI, ctypes.POINTER(PB), ctypes.c_int,
                                             ctypes.POINTER(PD), D, ctypes.c_int, PD]),
        "ref_amr_advance_pf": (ctypes.c_int, [PP, PF, ctypes.c_int, PG, PI, ctypes.POINTER(PB), ctypes.c_int,
                                              ctypes.POINTER(PD), ctypes.c_int, D, ctypes.c_int, ctypes.c_int, PD]),
        "ref_amr_advance": (ctypes.c_int, [PP, ctypes.c_int, PG, PI, ctypes.POINTER(PB), ctypes.c_int,
                                           ctypes.POINTER(PD), ctypes.c_int, D, ctypes.c_int, ctypes.c_int, PD]),
        "ref_level_tag": (ctypes.c_int, [PP, ctypes.c_int, PB, ctypes.c_int, PD,
                                         ctypes.POINTER(cmp_tag_criterion), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_ubyte)]),
        "ref_level_pf_mechanism": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, PD, D, D, D, D, PD, PD, PI]),
        "ref_pf_cac_mech": (D, [ctypes.c_int, PD, D, PD, PD, PD, PD, PD, D, D, D]),
        "ref_level_pf_gradq_implicit": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, PD, PD, D, PD, PI]),
        "ref_level_pf_gradq": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, PD, PD, D, PD]),
        "ref_level_pf_gradq_step": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, ctypes.c_int, PD, D,
                                                   PD, PD, PD]),
        "ref_level_pf_evolve": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, ctypes.c_int, PD, D, PD]),
        "ref_level_advance_pf": (ctypes.c_int, [PP, PF, PG, ctypes.c_int, PB, ctypes.c_int, PD,
                                                ctypes.c_int, D, ctypes.c_int, PD]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


# ---- pointwise wrappers --------------------------------------------------------------------------

def cons2prim(params, U):
    lib = load()
    U = np.ascontiguousarray(U, dtype=np.float64)
    W = np.empty_like(U)
    err = ctypes.c_int(0)
    lib.ref_cons2prim(ctypes.byref(params), _dptr(U), _dptr(W), ctypes.byref(err))
    return W


def prim2cons(params, W):
    lib = load()
    W = np.ascontiguousarray(W, dtype=np.float64)
    U = np.empty_like(W)
    err = ctypes.c_int(0)
    lib.ref_prim2cons(ctypes.byref(params), _dptr(W), _dptr(U), ctypes.byref(err))
    return U


def sound_speed(params, p, a1, a1r1, a2r2, id_k):
    err = ctypes.c_int(0)
    return load().ref_sound_speed(ctypes.byref(params), p, a1, a1r1, a2r2, id_k, ctypes.byref(err))


def nonconservative_K(params, p, a1, a1r1, a2r2, id_k):
    K = np.zeros(2)
    err = ctypes.c_int(0)
    load().ref_nonconservative_K(ctypes.byref(params), p, a1, a1r1, a2r2, id_k, _dptr(K), ctypes.byref(err))
    return K


def cons2flux(params, U, direction):
    U = np.ascontiguousarray(U, dtype=np.float64)
    F = np.empty_like(U)
    err = ctypes.c_int(0)
    load().ref_cons2flux(ctypes.byref(params), _dptr(U), direction, _dptr(F), ctypes.byref(err))
    return F


def weno5(v, face):
    return load().ref_weno5(float(v[0]), float(v[1]), float(v[2]), float(v[3]), float(v[4]), int(face))


def weno5_weights(v, face):
    f = np.zeros(3)
    w = np.zeros(3)
    load().ref_weno5_weights(*[float(x) for x in v], int(face), _dptr(f), _dptr(w))
    return f, w


def muscl(fm, fc, fp, dx, face, limiter):
    return load().ref_muscl(float(fm), float(fc), float(fp), float(dx), int(face), int(limiter))


REF_NMAX = 10   # row length of the oracle's stencil arrays (cpu_ref.h)


def weno5b_side(params, W5, face):
    """WENO5B primitive face state of the middle cell of W5 (5 x nstate primitives)."""
    W = np.zeros((5, REF_NMAX))
    W[:, :params.nstate] = np.asarray(W5, dtype=np.float64)
    out = np.zeros(REF_NMAX)
    load().ref_weno5b_side(ctypes.byref(params), _dptr(W), int(face), _dptr(out))
    return out[:params.nstate]


def thinc(beta, sigma, alpha_bar, face):
    return load().ref_thinc(float(beta), float(sigma), float(alpha_bar), int(face))


def riemann(params, UL, UR, direction):
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    F = np.zeros(params.nstate)
    vel = np.zeros(3)
    cmax = ctypes.c_double(0.0)
    err = ctypes.c_int(0)
    fn = load().ref_np_riemann if params.nphase else load().ref_riemann
    fn(ctypes.byref(params), _dptr(UL), _dptr(UR), direction, _dptr(F), _dptr(vel), ctypes.byref(cmax),
       ctypes.byref(err))
    return F, vel[:params.dim].copy(), cmax.value


# ---- the N-phase model (params.nphase = 2 or 3) ----------------------------------------------------

def np_eos_mixture(params, value, alpha, arho, output_internal_energy):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    r = np.ascontiguousarray(arho, dtype=np.float64)
    err = ctypes.c_int(0)
    return load().ref_np_eos_mixture(ctypes.byref(params), float(value), _dptr(a), _dptr(r),
                                     int(output_internal_energy), ctypes.byref(err))


def np_cons2prim(params, U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    W = np.empty_like(U)
    err = ctypes.c_int(0)
    load().ref_np_cons2prim(ctypes.byref(params), _dptr(U), _dptr(W), ctypes.byref(err))
    return W


def np_prim2cons(params, W):
    W = np.ascontiguousarray(W, dtype=np.float64)
    U = np.empty_like(W)
    err = ctypes.c_int(0)
    load().ref_np_prim2cons(ctypes.byref(params), _dptr(W), _dptr(U), ctypes.byref(err))
    return U


def np_sound_speed(params, p, alpha, arho, id_k):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    r = np.ascontiguousarray(arho, dtype=np.float64)
    err = ctypes.c_int(0)
    return load().ref_np_sound_speed(ctypes.byref(params), float(p), _dptr(a), _dptr(r), int(id_k),
                                     ctypes.byref(err))


def np_K(params, p, alpha, arho, id_k):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    r = np.ascontiguousarray(arho, dtype=np.float64)
    K = np.zeros(params.nphase)
    err = ctypes.c_int(0)
    load().ref_np_K(ctypes.byref(params), float(p), _dptr(a), _dptr(r), int(id_k), _dptr(K),
                    ctypes.byref(err))
    return K


def np_cons2flux(params, U, direction):
    U = np.ascontiguousarray(U, dtype=np.float64)
    F = np.empty_like(U)
    err = ctypes.c_int(0)
    load().ref_np_cons2flux(ctypes.byref(params), _dptr(U), int(direction), _dptr(F), ctypes.byref(err))
    return F


def np_limiter_consistent(slopes):
    s = np.ascontiguousarray(slopes, dtype=np.float64)
    th = np.zeros(len(s))
    load().ref_np_limiter_consistent(len(s), _dptr(s), _dptr(th))
    return th


def np_limit_consistent(params, Uface, Ucell, dx, face):
    Uf = np.ascontiguousarray(Uface, dtype=np.float64)
    Uc = np.ascontiguousarray(Ucell, dtype=np.float64)
    out = np.empty_like(Uf)
    err = ctypes.c_int(0)
    load().ref_np_limit_consistent(ctypes.byref(params), _dptr(Uf), _dptr(Uc), float(dx), int(face),
                                   _dptr(out), ctypes.byref(err))
    return out


# ---- box and level wrappers ----------------------------------------------------------------------

def rhs_box(params, geom, box, ng, Ufab, want_faces=False):
    """dUdt over the valid box [n, (k,) j, i], cmax, and optionally the face arrays."""
    from compas_amd.capi import cmp_box
    lib = load()
    dim = params.dim
    b = cmp_box()
    lo, hi = box
    for d in range(3):
        b.lo[d] = lo[d] if d < dim else 0
        b.hi[d] = hi[d] if d < dim else 0
    nv = [hi[d] - lo[d] + 1 for d in range(dim)]
    N = params.nstate
    dUdt = np.zeros((N,) + tuple(reversed(nv)))
    cmax = ctypes.c_double(0.0)
    F = (ctypes.POINTER(ctypes.c_double) * 3)()
    FN = (ctypes.POINTER(ctypes.c_double) * 3)()
    faces, facesnc = [], []
    if want_faces:
        for d in range(dim):
            fe = list(nv)
            fe[d] += 1
            a = np.zeros((N,) + tuple(reversed(fe)))
            an = np.zeros(tuple(reversed(fe)))
            faces.append(a)
            facesnc.append(an)
            F[d] = _dptr(a)
            FN[d] = _dptr(an)
    Ufab = np.ascontiguousarray(Ufab, dtype=np.float64)
    rc = lib.ref_rhs_box(ctypes.byref(params), ctypes.byref(geom), ctypes.byref(b), ng, _dptr(Ufab),
                         _dptr(dUdt), ctypes.byref(cmax), F if want_faces else None,
                         FN if want_faces else None)
    if rc:
        raise RuntimeError("ref_rhs_box failed: %d" % rc)
    if want_faces:
        return dUdt, cmax.value, faces, facesnc
    return dUdt, cmax.value


def level_fill_ghosts(params, geom, layout, buf):
    rc = load().ref_level_fill_ghosts(ctypes.byref(params), ctypes.byref(geom), layout.nbox,
                                      layout.cboxes, layout.ng, _dptr(buf))
    if rc:
        raise RuntimeError("ref_level_fill_ghosts failed: %d" % rc)


def level_rhs(params, geom, layout, buf):
    dUdt = np.zeros_like(buf)
    cmax = ctypes.c_double(0.0)
    rc = load().ref_level_rhs(ctypes.byref(params), ctypes.byref(geom), layout.nbox, layout.cboxes,
                              layout.ng, _dptr(buf), _dptr(dUdt), ctypes.byref(cmax))
    if rc:
        raise RuntimeError("ref_level_rhs failed: %d" % rc)
    return dUdt, cmax.value


def level_rk3_step(params, geom, layout, buf, dt):
    cmax = ctypes.c_double(0.0)
    rc = load().ref_level_rk3_step(ctypes.byref(params), ctypes.byref(geom), layout.nbox,
                                   layout.cboxes, layout.ng, _dptr(buf), float(dt), ctypes.byref(cmax))
    if rc:
        raise RuntimeError("ref_level_rk3_step failed: %d" % rc)
    return cmax.value


def level_initial_cmax(params, geom, layout, buf):
    cmax = ctypes.c_double(0.0)
    rc = load().ref_level_initial_cmax(ctypes.byref(params), ctypes.byref(geom), layout.nbox,
                                       layout.cboxes, layout.ng, _dptr(buf), ctypes.byref(cmax))
    if rc:
        raise RuntimeError("ref_level_initial_cmax failed: %d" % rc)
    return cmax.value


def compute_dt(geom, cmax, cfl, limiter, dt_prev):
    return load().ref_compute_dt(ctypes.byref(geom), float(cmax), float(cfl), int(limiter), float(dt_prev))


def level_advance(params, geom, layout, buf, nsteps, cfl, limiter=1):
    hist = np.zeros(max(1, nsteps))
    rc = load().ref_level_advance(ctypes.byref(params), ctypes.byref(geom), layout.nbox,
                                  layout.cboxes, layout.ng, _dptr(buf), int(nsteps), float(cfl),
                                  int(limiter), _dptr(hist))
    if rc:
        raise RuntimeError("ref_level_advance failed: %d" % rc)
    return hist[:nsteps]


# ---- phase field ---------------------------------------------------------------------------------

def pf_map_backward(phi, eta):
    return load().ref_pf_map_backward(float(phi), float(eta))


def pf_map_forward(psi, eta, deriv=0):
    return load().ref_pf_map_forward(float(psi), float(eta), int(deriv))


def pf_acdi_flux(pf_m, pf_c, n_m, n_c, mob, eta, dx):
    return load().ref_pf_acdi_flux(*[float(x) for x in (pf_m, pf_c, n_m, n_c, mob, eta, dx)])


def pf_gradq_face(pf_m, pf_c, jhat, eta, eta_factor):
    return load().ref_pf_gradq_face(*[float(x) for x in (pf_m, pf_c, jhat, eta, eta_factor)])


def pf_flux(params, pf, UL, UR, g):
    """fhat (nstate) and the UD wave speed of one face."""
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    F = np.zeros(params.nstate)
    c = ctypes.c_double(0.0)
    err = ctypes.c_int(0)
    load().ref_pf_flux(ctypes.byref(params), ctypes.byref(pf), _dptr(UL), _dptr(UR), float(g),
                       _dptr(F), ctypes.byref(c), ctypes.byref(err))
    return F, c.value


def np_cons2phase(params, pf, U):
    """Phasic conservative vector Cp[q*nphase + k] (q: mass, momenta, energy, alpha) of one state."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    Cp = np.zeros((3 + params.dim) * params.nphase)
    err = ctypes.c_int(0)
    load().ref_np_cons2phase(ctypes.byref(params), ctypes.byref(pf), _dptr(U), _dptr(Cp), ctypes.byref(err))
    return Cp


def np_pf_flux(params, pf, UL, UR, g):
    """N-phase model: fhat (nstate) and the UD wave speed of one face; g holds Grad_Q per phase."""
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    gg = np.ascontiguousarray(g, dtype=np.float64)
    F = np.zeros(params.nstate)
    c = ctypes.c_double(0.0)
    err = ctypes.c_int(0)
    load().ref_np_pf_flux(ctypes.byref(params), ctypes.byref(pf), _dptr(UL), _dptr(UR), _dptr(gg), _dptr(F),
                          ctypes.byref(c), ctypes.byref(err))
    return F, c.value


def pf_coefficients(pf, geom, max_v):
    L, e, m = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    load().ref_pf_coefficients(ctypes.byref(pf), ctypes.byref(geom), float(max_v), ctypes.byref(L),
                               ctypes.byref(e), ctypes.byref(m))
    return L.value, e.value, m.value


def pf_compute_dt(pf, geom, gradq_max, pf_time, dt, prev):
    p = ctypes.c_double(prev)
    r = load().ref_pf_compute_dt(ctypes.byref(pf), ctypes.byref(geom), float(gradq_max), float(pf_time),
                                 float(dt), ctypes.byref(p))
    return r, p.value


def level_fill_pf(geom, layout, buf):
    """Ghost fill of a Phase-Field field level buffer (periodic or foextrap)."""
    rc = load().ref_level_fill_pf(layout.dim, ctypes.byref(geom), layout.nbox, layout.cboxes,
                                  layout.ng, layout.ncomp, _dptr(buf))
    if rc:
        raise RuntimeError("ref_level_fill_pf failed: %d" % rc)


def level_pf_rhs(params, pf, geom, layout, buf, gradq):
    dUdt = np.zeros_like(buf)
    rc = load().ref_level_pf_rhs(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom), layout.nbox,
                                 layout.cboxes, layout.ng, _dptr(buf), _dptr(gradq), _dptr(dUdt))
    if rc:
        raise RuntimeError("ref_level_pf_rhs failed: %d" % rc)
    return dUdt


def level_pf_rk3_step(params, pf, geom, layout, buf, gradq, pf_dt):
    rc = load().ref_level_pf_rk3_step(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom),
                                      layout.nbox, layout.cboxes, layout.ng, _dptr(buf), _dptr(gradq),
                                      float(pf_dt))
    if rc:
        raise RuntimeError("ref_level_pf_rk3_step failed: %d" % rc)


def level_average_down(params, flayout, fbuf, clayout, cbuf, ratio):
    """In place on cbuf: the coarse cells under the fine boxes take their children's mean."""
    rc = load().ref_level_average_down(ctypes.byref(params), flayout.nbox, flayout.cboxes, flayout.ng, _dptr(fbuf),
                                       clayout.nbox, clayout.cboxes, clayout.ng, _dptr(cbuf), int(ratio))
    if rc:
        raise RuntimeError("ref_level_average_down failed: %d" % rc)


def level_fill_two_levels(params, fgeom, flayout, fbuf, cgeom, clayout, cbuf, ratio):
    """In place on fbuf: every ghost cell of the fine level (fine sources copied, coarse-fine ghosts
    by lincc from cbuf)."""
    rc = load().ref_level_fill_two_levels(ctypes.byref(params), ctypes.byref(fgeom), flayout.nbox, flayout.cboxes,
                                          flayout.ng, _dptr(fbuf), ctypes.byref(cgeom), clayout.nbox, clayout.cboxes,
                                          clayout.ng, _dptr(cbuf), int(ratio))
    if rc:
        raise RuntimeError("ref_level_fill_two_levels failed: %d" % rc)


def level_rk_step(params, geom, layout, buf, dt, integrator):
    """One step of size dt in place on buf: integrator 1 ForwardEuler, 2 TVD-RK2, 3 TVD-RK3, 4 RK4.
    Returns the c_max of the step's stages."""
    cm = np.zeros(1)
    rc = load().ref_level_rk_step(ctypes.byref(params), ctypes.byref(geom), layout.nbox, layout.cboxes, layout.ng,
                                  _dptr(buf), float(dt), _dptr(cm), int(integrator))
    if rc:
        raise RuntimeError("ref_level_rk_step failed: %d" % rc)
    return cm[0]


def level_pf_rk_step(params, pf, geom, layout, buf, gradq, pf_dt, integrator):
    """One Phase-Field sub-step in place on buf at the fixed face buffer gradq."""
    rc = load().ref_level_pf_rk_step(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom), layout.nbox,
                                     layout.cboxes, layout.ng, _dptr(buf), _dptr(gradq), float(pf_dt), int(integrator))
    if rc:
        raise RuntimeError("ref_level_pf_rk_step failed: %d" % rc)


def level_advance_int(params, geom, layout, buf, nsteps, cfl, integrator, pf=None, limiter=1):
    """nsteps steps (with the Phase-Field step when pf is given) with run.TimeIntegrator; returns
    the dt history."""
    hist = np.zeros(max(1, nsteps))
    lib = load()
    if pf is None:
        rc = lib.ref_level_advance_int(ctypes.byref(params), ctypes.byref(geom), layout.nbox, layout.cboxes,
                                       layout.ng, _dptr(buf), int(nsteps), float(cfl), int(limiter), _dptr(hist),
                                       int(integrator))
    else:
        rc = lib.ref_level_advance_pf_int(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom), layout.nbox,
                                          layout.cboxes, layout.ng, _dptr(buf), int(nsteps), float(cfl), int(limiter),
                                          _dptr(hist), int(integrator))
    if rc:
        raise RuntimeError("ref_level_advance_int failed: %d" % rc)
    return hist[:nsteps]


def level_fill_from(params, fgeom, new_layout, new_buf, old_layout, old_buf, cgeom, clayout, cbuf, ratio):
    """In place on new_buf: every cell of the new boxes from the old boxes of the level (old_layout
    None: none), else by interpolation from the coarse level."""
    from compas_amd.capi import cmp_box
    nbo = old_layout.nbox if old_layout is not None else 0
    ob = old_layout.cboxes if old_layout is not None else (cmp_box * 1)()
    ngo = old_layout.ng if old_layout is not None else 0
    obuf = old_buf if old_buf is not None else np.zeros(1)
    rc = load().ref_level_fill_from(ctypes.byref(params), ctypes.byref(fgeom), new_layout.nbox, new_layout.cboxes,
                                    new_layout.ng, _dptr(new_buf), nbo, ob, ngo, _dptr(obuf), ctypes.byref(cgeom),
                                    clayout.nbox, clayout.cboxes, clayout.ng, _dptr(cbuf), int(ratio))
    if rc:
        raise RuntimeError("ref_level_fill_from failed: %d" % rc)


def level_tag(params, layout, buf, criteria):
    """Tags of every valid cell (bytes, box by box), buf ghost-filled; criteria as
    compas_amd.tag_criteria takes them."""
    from compas_amd import tag_criteria
    arr = tag_criteria(criteria)
    n = sum(int(np.prod([h - l + 1 for l, h in zip(lo, hi)])) for lo, hi in layout.boxes)
    tags = np.zeros(n, dtype=np.uint8)
    rc = load().ref_level_tag(ctypes.byref(params), layout.nbox, layout.cboxes, layout.ng, _dptr(buf), arr,
                              len(criteria), tags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    if rc:
        raise RuntimeError("ref_level_tag failed: %d" % rc)
    return tags


def level_max_velocity(params, layout, buf):
    return load().ref_level_max_velocity(ctypes.byref(params), layout.nbox, layout.cboxes, layout.ng,
                                         _dptr(buf))


def level_pf_mechanism(params, pf, geom, alayout, alpha, dt, mob, eta, lam, nfaces, implicit=False):
    """Jhat (face buffer of nfaces), or Div_Jhat (alpha layout) with implicit=True, and the number
    of sub-steps; alpha: 1 comp, 1 ghost, filled."""
    jhat = None if implicit else np.zeros(nfaces)
    dj = np.zeros_like(alpha) if implicit else None
    n = ctypes.c_int(0)
    rc = load().ref_level_pf_mechanism(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom),
                                       alayout.nbox, alayout.cboxes, _dptr(alpha), float(dt), float(mob),
                                       float(eta), float(lam), _dptr(jhat), _dptr(dj), ctypes.byref(n))
    if rc:
        raise RuntimeError("ref_level_pf_mechanism failed: %d" % rc)
    return (dj if implicit else jhat), n.value


def pf_cac_mech(dim, dx, f, fm, fp, nc, nm, np_, mob, lam, eta):
    """CAC-Adv mechanism of one cell; dx, fm, fp, nc, nm, np_: sequences of dim values."""
    def v3(x):
        a = np.zeros(3)
        a[:dim] = np.asarray(x, dtype=np.float64)[:dim]
        return a
    arr = [v3(x) for x in (dx, fm, fp, nc, nm, np_)]
    return load().ref_pf_cac_mech(int(dim), _dptr(arr[0]), float(f), _dptr(arr[1]), _dptr(arr[2]),
                                  _dptr(arr[3]), _dptr(arr[4]), _dptr(arr[5]), float(mob), float(lam),
                                  float(eta))


def level_pf_gradq_implicit(params, pf, geom, alayout, alpha, divjhat, eta, nfaces):
    """(Grad_Q face buffer, CG iterations) of the implicit RC."""
    gq = np.zeros(nfaces)
    it = ctypes.c_int(0)
    rc = load().ref_level_pf_gradq_implicit(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom),
                                            alayout.nbox, alayout.cboxes, _dptr(alpha), _dptr(divjhat),
                                            float(eta), _dptr(gq), ctypes.byref(it))
    if rc:
        raise RuntimeError("ref_level_pf_gradq_implicit failed: %d" % rc)
    return gq, it.value


def level_pf_gradq(params, pf, geom, alayout, alpha, jhat, eta):
    gq = np.zeros_like(jhat)
    rc = load().ref_level_pf_gradq(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom),
                                   alayout.nbox, alayout.cboxes, _dptr(alpha), _dptr(jhat), float(eta),
                                   _dptr(gq))
    if rc:
        raise RuntimeError("ref_level_pf_gradq failed: %d" % rc)
    return gq


def level_pf_gradq_step(params, pf, geom, layout, buf, dt, nfaces):
    """(gradq face buffer, max|Grad_Q|, [max|u|, mobility, eta, mechanism sub-steps])."""
    gq = np.zeros(nfaces)
    gm = ctypes.c_double(0.0)
    info = np.zeros(4)
    rc = load().ref_level_pf_gradq_step(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom),
                                        layout.nbox, layout.cboxes, layout.ng, _dptr(buf), float(dt),
                                        _dptr(gq), ctypes.byref(gm), _dptr(info))
    if rc:
        raise RuntimeError("ref_level_pf_gradq_step failed: %d" % rc)
    return gq, gm.value, list(info)


def level_pf_evolve(params, pf, geom, layout, buf, dt):
    """In place on buf; returns [max|u|, mobility, eta, mech sub-steps, max|Grad_Q|, PF sub-steps]."""
    info = np.zeros(6)
    rc = load().ref_level_pf_evolve(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom),
                                    layout.nbox, layout.cboxes, layout.ng, _dptr(buf), float(dt),
                                    _dptr(info))
    if rc:
        raise RuntimeError("ref_level_pf_evolve failed: %d" % rc)
    return list(info)


def level_advance_pf(params, pf, geom, layout, buf, nsteps, cfl, limiter=1):
    hist = np.zeros(max(1, nsteps))
    rc = load().ref_level_advance_pf(ctypes.byref(params), ctypes.byref(pf), ctypes.byref(geom),
                                     layout.nbox, layout.cboxes, layout.ng, _dptr(buf), int(nsteps),
                                     float(cfl), int(limiter), _dptr(hist))
    if rc:
        raise RuntimeError("ref_level_advance_pf failed: %d" % rc)
    return hist[:nsteps]


def _amr_args(geoms, layouts, bufs):
    from compas_amd.capi import cmp_box, cmp_geometry
    n = len(layouts)
    if len({l.ng for l in layouts}) != 1:
        raise ValueError("the levels of a hierarchy share their ghost width")
    G = (cmp_geometry * n)(*geoms)
    nb = (ctypes.c_int * n)(*[l.nbox for l in layouts])
    B = (ctypes.POINTER(cmp_box) * n)(*[ctypes.cast(l.cboxes, ctypes.POINTER(cmp_box)) for l in layouts])
    U = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(b) for b in bufs])
    return n, G, nb, B, U


def amr_rk3_step(params, geoms, layouts, bufs, dt, ratio):
    """One stage-coupled RK3 step of all levels (coarsest first), in place on bufs; returns the
    levels' c_max."""
    n, G, nb, B, U = _amr_args(geoms, layouts, bufs)
    cmax = np.zeros(n)
    rc = load().ref_amr_rk3_step(ctypes.byref(params), n, G, nb, B, layouts[0].ng, U, float(dt), int(ratio),
                                 _dptr(cmax))
    if rc:
        raise RuntimeError("ref_amr_rk3_step failed: %d" % rc)
    return cmax


def amr_compute_dt(geoms, cmax, cfl, limiter, dt_prev):
    from compas_amd.capi import cmp_geometry
    n = len(geoms)
    G = (cmp_geometry * n)(*geoms)
    c = np.ascontiguousarray(cmax, dtype=np.float64)
    return load().ref_amr_compute_dt(n, G, _dptr(c), float(cfl), int(limiter), float(dt_prev))


def amr_advance(params, geoms, layouts, bufs, nsteps, cfl, ratio, limiter=1):
    """nsteps steps of the hierarchy in place on bufs; returns the dt of every step."""
    n, G, nb, B, U = _amr_args(geoms, layouts, bufs)
    hist = np.zeros(max(1, nsteps))
    rc = load().ref_amr_advance(ctypes.byref(params), n, G, nb, B, layouts[0].ng, U, int(nsteps), float(cfl),
                                int(limiter), int(ratio), _dptr(hist))
    if rc:
        raise RuntimeError("ref_amr_advance failed: %d" % rc)
    return hist[:nsteps]


def amr_pf_gradq_step(params, pf, geoms, layouts, bufs, dt, ratio):
    """Grad_Q of every level of a hierarchy for a flow step of size dt (PhaseField_Evolve's
    mechanism and Grad_Q over the levels): ([face buffer per level, cmp_face_layout with one
    component per volume fraction], [max|u|, mobility, eta, mechanism sub-steps, CG iterations])."""
    from compas_amd import FaceLayout
    n, G, nb, B, U = _amr_args(geoms, layouts, bufs)
    npf = params.nphase if params.nphase else 1
    gqs = [np.zeros(FaceLayout(params.dim, l.boxes, npf).total) for l in layouts]
    GQ = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(g) for g in gqs])
    info = np.zeros(5)
    rc = load().ref_amr_pf_gradq_step(ctypes.byref(params), ctypes.byref(pf), n, G, nb, B, layouts[0].ng, U,
                                      float(dt), int(ratio), GQ, _dptr(info))
    if rc:
        raise RuntimeError("ref_amr_pf_gradq_step failed: %d" % rc)
    return gqs, list(info)


def amr_pf_gradq_implicit(params, pf, geoms, alayouts, alphas, rhss, eta, ratio):
    """The composite implicit solve of a two-phase hierarchy: alphas and rhss are 1-comp, 1-ghost
    level buffers (alpha's ghosts are filled in place); returns ([Grad_Q per level], [b per level],
    CG iterations)."""
    from compas_amd import FaceLayout
    n, G, nb, B, A = _amr_args(geoms, alayouts, alphas)
    R = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(r) for r in rhss])
    gqs = [np.zeros(FaceLayout(params.dim, l.boxes).total) for l in alayouts]
    bfs = [np.zeros(FaceLayout(params.dim, l.boxes).total) for l in alayouts]
    GQ = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(g) for g in gqs])
    BF = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(g) for g in bfs])
    it = ctypes.c_int(0)
    rc = load().ref_amr_pf_gradq_implicit(ctypes.byref(params), ctypes.byref(pf), n, G, nb, B, A, R, float(eta),
                                          int(ratio), GQ, BF, ctypes.byref(it))
    if rc:
        raise RuntimeError("ref_amr_pf_gradq_implicit failed: %d" % rc)
    return gqs, bfs, it.value


def amr_pf_evolve(params, pf, geoms, layouts, bufs, dt, ratio):
    """The Phase-Field step of a hierarchy, in place on bufs; returns [max|u|, mobility, eta, mechanism
    sub-steps, max|Grad_Q|, Phase-Field sub-steps, CG iterations]."""
    n, G, nb, B, U = _amr_args(geoms, layouts, bufs)
    info = np.zeros(7)
    rc = load().ref_amr_pf_evolve(ctypes.byref(params), ctypes.byref(pf), n, G, nb, B, layouts[0].ng, U, float(dt),
                                  int(ratio), _dptr(info))
    if rc:
        raise RuntimeError("ref_amr_pf_evolve failed: %d" % rc)
    return list(info)


def amr_advance_pf(params, pf, geoms, layouts, bufs, nsteps, cfl, ratio, limiter=1):
    """nsteps flow + Phase-Field steps of a hierarchy in place on bufs; returns every step's dt."""
    n, G, nb, B, U = _amr_args(geoms, layouts, bufs)
    hist = np.zeros(max(1, nsteps))
    rc = load().ref_amr_advance_pf(ctypes.byref(params), ctypes.byref(pf), n, G, nb, B, layouts[0].ng, U,
                                   int(nsteps), float(cfl), int(limiter), int(ratio), _dptr(hist))
    if rc:
        raise RuntimeError("ref_amr_advance_pf failed: %d" % rc)
    return hist[:nsteps]


def set_num_threads(n):
    load().ref_set_num_threads(int(n))


def num_threads():
    return load().ref_num_threads()
