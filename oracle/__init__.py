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


def _dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        from . import build as _b
        _b.build()
    # the product library's struct mirrors are shared (plain PODs)
    from compas_amd.capi import cmp_box, cmp_geometry, cmp_params
    lib = ctypes.CDLL(LIB_PATH)
    D = ctypes.c_double
    PD = ctypes.POINTER(D)
    PI = ctypes.POINTER(ctypes.c_int)
    PP = ctypes.POINTER(cmp_params)
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


def weno5b_side(params, W5, face):
    """WENO5B primitive face state of the middle cell of W5 (5 x nstate primitives)."""
    W = np.zeros((5, 7))
    W[:, :params.nstate] = np.asarray(W5, dtype=np.float64)
    out = np.zeros(7)
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
    load().ref_riemann(ctypes.byref(params), _dptr(UL), _dptr(UR), direction, _dptr(F), _dptr(vel),
                       ctypes.byref(cmax), ctypes.byref(err))
    return F, vel[:params.dim].copy(), cmax.value


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


def set_num_threads(n):
    load().ref_set_num_threads(int(n))


def num_threads():
    return load().ref_num_threads()
