"""ctypes mirror of include/compas_amd.h.

The shared library is the product: this module only declares its structs and entry points.
Loading fails loudly when libcompas_amd.so has not been built; there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcompas_amd.so")

CMP_OK = 0
CMP_ERR_ARG = 1
CMP_ERR_UNSUPPORTED = 2
CMP_ERR_HIP = 3
CMP_ERR_NAN = 4
CMP_ERR_COVOLUME = 5
CMP_ERR_BC = 6
CMP_ERR_LAYOUT = 7
CMP_ERR_NOMEM = 8
CMP_ERR_SCRATCH = 9
CMP_ERR_DENSITY = 10
CMP_ERR_PRESSURE = 11
CMP_ERR_ALPHA = 12
CMP_NOALLOC_MAX_BOXES = 4096
CMP_NOALLOC_MAX_TILES = 65536

ERROR_NAMES = {
    CMP_OK: "CMP_OK",
    CMP_ERR_ARG: "CMP_ERR_ARG",
    CMP_ERR_UNSUPPORTED: "CMP_ERR_UNSUPPORTED",
    CMP_ERR_HIP: "CMP_ERR_HIP",
    CMP_ERR_NAN: "CMP_ERR_NAN",
    CMP_ERR_COVOLUME: "CMP_ERR_COVOLUME",
    CMP_ERR_BC: "CMP_ERR_BC",
    CMP_ERR_LAYOUT: "CMP_ERR_LAYOUT",
    CMP_ERR_NOMEM: "CMP_ERR_NOMEM",
    CMP_ERR_SCRATCH: "CMP_ERR_SCRATCH",
    CMP_ERR_DENSITY: "CMP_ERR_DENSITY",
    CMP_ERR_PRESSURE: "CMP_ERR_PRESSURE",
    CMP_ERR_ALPHA: "CMP_ERR_ALPHA",
}


class CmpError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__("%s failed: %s (%d)" % (where, ERROR_NAMES.get(code, "?"), code))


class cmp_params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("nstate", ctypes.c_int),
        ("nphase", ctypes.c_int),
        ("eos_A", ctypes.c_double * 3),
        ("eos_B", ctypes.c_double * 3),
        ("eos_C", ctypes.c_double * 3),
        ("eos_D", ctypes.c_double * 3),
        ("id_weno", ctypes.c_int),
        ("id_limiter", ctypes.c_int),
        ("id_poly", ctypes.c_int),
        ("id_bound", ctypes.c_int),
        ("id_thinc", ctypes.c_int),
        ("thinc_use_normal", ctypes.c_int),
        ("thinc_eps", ctypes.c_double),
        ("thinc_beta", ctypes.c_double),
        ("thinc_normal_threshold", ctypes.c_double),
        ("source_term", ctypes.c_int),
        ("id_riemann", ctypes.c_int),
    ]


class cmp_geometry(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("domain_lo", ctypes.c_int * 3),
        ("domain_hi", ctypes.c_int * 3),
        ("is_periodic", ctypes.c_int * 3),
        ("lo_bc", ctypes.c_int * 3),
        ("hi_bc", ctypes.c_int * 3),
        ("dx", ctypes.c_double * 3),
        ("prob_lo", ctypes.c_double * 3),
    ]


class cmp_box(ctypes.Structure):
    _fields_ = [
        ("lo", ctypes.c_int * 3),
        ("hi", ctypes.c_int * 3),
        ("offset", ctypes.c_longlong),
    ]


class cmp_pf_params(ctypes.Structure):
    _fields_ = [
        ("mechanism", ctypes.c_int),
        ("explicit_rc", ctypes.c_int),
        ("id_flux", ctypes.c_int),
        ("flux_phase2flux", ctypes.c_int),
        ("cfl", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("eta_factor", ctypes.c_double),
        ("eta_multiplier", ctypes.c_double),
        ("p_min", ctypes.c_double * 3),
        ("max_step", ctypes.c_int),
        ("change_limiter", ctypes.c_int),
    ]


class cmp_tag_criterion(ctypes.Structure):
    _fields_ = [("var", ctypes.c_int), ("kind", ctypes.c_int), ("v0", ctypes.c_double), ("v1", ctypes.c_double)]


CMP_TAG_RHO, CMP_TAG_P, CMP_TAG_U, CMP_TAG_V, CMP_TAG_W = -1, -2, -3, -4, -5
CMP_TAG_GRAD, CMP_TAG_LESS, CMP_TAG_GREATER, CMP_TAG_RANGE = 0, 1, 2, 3


P = ctypes.POINTER
vp = ctypes.c_void_p
dp = ctypes.c_void_p   # device pointers travel as integers (torch data_ptr())

# name -> (restype, argtypes); exactly the entry points include/compas_amd.h declares
SIGNATURES = {
    "cmp_version": (ctypes.c_char_p, []),
    "cmp_abi_sizes": (ctypes.c_int, [P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]),
    "cmp_fab_size": (ctypes.c_longlong, [ctypes.c_int, P(cmp_box), ctypes.c_int, ctypes.c_int]),
    "cmp_level_layout": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(cmp_box), ctypes.c_int,
                                        ctypes.c_int, P(ctypes.c_longlong)]),
    "cmp_rhs_box": (ctypes.c_int, [P(cmp_params), P(cmp_geometry), P(cmp_box), ctypes.c_int,
                                   dp, dp, dp, dp, dp, dp, dp, dp, dp, vp]),
    "cmp_cmax_box": (ctypes.c_int, [P(cmp_params), P(cmp_geometry), P(cmp_box), ctypes.c_int, dp, dp, vp]),
    "cmp_pf_rhs_box": (ctypes.c_int, [P(cmp_params), P(cmp_pf_params), P(cmp_geometry), P(cmp_box), ctypes.c_int,
                                      dp, dp, dp, vp]),
    "cmp_level_create": (ctypes.c_int, [P(cmp_params), P(cmp_geometry), ctypes.c_int, P(cmp_box),
                                        ctypes.c_int, P(vp)]),
    "cmp_level_destroy": (ctypes.c_int, [vp]),
    "cmp_level_fill_ghosts": (ctypes.c_int, [vp, dp, vp]),
    "cmp_level_rhs": (ctypes.c_int, [vp, dp, dp, dp, vp]),
    "cmp_level_rk3_stage": (ctypes.c_int, [vp, ctypes.c_int, dp, dp, dp, dp, dp, vp]),
    "cmp_level_rk3_step": (ctypes.c_int, [vp, dp, dp, dp, dp, dp, vp]),
    "cmp_level_initial_cmax": (ctypes.c_int, [vp, dp, dp, vp]),
    "cmp_level_compute_dt": (ctypes.c_int, [vp, dp, dp, ctypes.c_double, ctypes.c_int, vp]),
    "cmp_level_advance": (ctypes.c_int, [vp, dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         dp, dp, dp, ctypes.c_int, ctypes.c_int, vp]),
    "cmp_level_check": (ctypes.c_int, [vp, vp]),
    "cmp_level_info": (ctypes.c_int, [vp, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                      P(ctypes.c_int)]),
    "cmp_level_set_tiles": (ctypes.c_int, [vp, P(ctypes.c_int), ctypes.c_int]),
    "cmp_level_tile_shapes": (ctypes.c_int, [vp, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]),
    "cmp_level_tune_tiles": (ctypes.c_int, [vp, dp, ctypes.c_int, P(ctypes.c_int), ctypes.c_int, vp]),
    "cmp_tile_tuning": (ctypes.c_int, [vp, P(ctypes.c_int), P(ctypes.c_double), P(ctypes.c_int), P(ctypes.c_int),
                                       P(ctypes.c_double)]),
    "cmp_tile_cache_save": (ctypes.c_int, [ctypes.c_char_p]),
    "cmp_tile_cache_load": (ctypes.c_int, [ctypes.c_char_p]),
    "cmp_tile_cache_clear": (ctypes.c_int, []),
    "cmp_cluster_blocks": (ctypes.c_int, [P(ctypes.c_ubyte), ctypes.c_int, P(ctypes.c_int), ctypes.c_double, ctypes.c_int,
                                          P(ctypes.c_int), P(ctypes.c_int)]),
    "cmp_face_layout": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(cmp_box), ctypes.c_int,
                                       P(ctypes.c_longlong), P(ctypes.c_longlong)]),
    "cmp_abi_pf_size": (ctypes.c_int, [P(ctypes.c_int)]),
    "cmp_level_pf_setup": (ctypes.c_int, [vp, P(cmp_pf_params)]),
    "cmp_level_pf_rhs": (ctypes.c_int, [vp, dp, dp, dp, vp]),
    "cmp_level_pf_gradq": (ctypes.c_int, [vp, dp, ctypes.c_double, dp, P(ctypes.c_double),
                                          P(ctypes.c_double), vp]),
    "cmp_level_pf_rk3_step": (ctypes.c_int, [vp, dp, dp, dp, dp, ctypes.c_double, vp]),
    "cmp_level_pf_solver_iterations": (ctypes.c_int, [vp, P(ctypes.c_int)]),
    "cmp_level_pf_evolve": (ctypes.c_int, [vp, dp, dp, dp, dp, P(ctypes.c_double), vp]),
    "cmp_level_tag": (ctypes.c_int, [vp, dp, P(cmp_tag_criterion), ctypes.c_int, dp, vp]),
    "cmp_level_average_down": (ctypes.c_int, [vp, dp, vp, dp, ctypes.c_int, vp]),
    "cmp_level_fill_two_levels": (ctypes.c_int, [vp, dp, vp, dp, ctypes.c_int, vp]),
    "cmp_level_set_integrator": (ctypes.c_int, [vp, ctypes.c_int]),
    "cmp_level_set_time": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, vp]),
    "cmp_level_get_time": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), vp]),
    "cmp_level_rk_step": (ctypes.c_int, [vp, ctypes.c_int, dp, dp, dp, dp, dp, dp, vp]),
    "cmp_level_pf_rk_step": (ctypes.c_int, [vp, ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_double, vp]),
    "cmp_level_fill_from": (ctypes.c_int, [vp, dp, vp, dp, vp, dp, ctypes.c_int, vp]),
    "cmp_amr_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]),
    "cmp_amr_destroy": (ctypes.c_int, [vp]),
    "cmp_amr_info": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "cmp_amr_tiles": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int)]),
    "cmp_amr_tile_info": (ctypes.c_int, [vp, P(ctypes.c_int), P(ctypes.c_int)]),
    "cmp_amr_initial_cmax": (ctypes.c_int, [vp, ctypes.POINTER(dp), dp, vp]),
    "cmp_amr_compute_dt": (ctypes.c_int, [vp, dp, dp, ctypes.c_double, ctypes.c_int, vp]),
    "cmp_amr_rk3_step": (ctypes.c_int, [vp, ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(dp), dp, dp,
                                        vp]),
    "cmp_amr_advance": (ctypes.c_int, [vp, ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, dp, dp, dp, ctypes.c_int, vp]),
    "cmp_amr_pf_setup": (ctypes.c_int, [vp, P(cmp_pf_params)]),
    "cmp_amr_pf_gradq": (ctypes.c_int, [vp, ctypes.POINTER(dp), ctypes.c_double, ctypes.POINTER(dp),
                                        P(ctypes.c_double), P(ctypes.c_double), vp]),
    "cmp_amr_pf_evolve": (ctypes.c_int, [vp, ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(dp), dp,
                                         P(ctypes.c_double), vp]),
    "cmp_amr_pf_solver_iterations": (ctypes.c_int, [vp, P(ctypes.c_int)]),
    "cmp_batch_create": (ctypes.c_int, [P(cmp_params), ctypes.c_longlong, ctypes.c_int, P(vp)]),
    "cmp_batch_destroy": (ctypes.c_int, [vp]),
    "cmp_batch_admit": (ctypes.c_int, [vp, P(cmp_params), P(cmp_geometry), ctypes.c_int, P(cmp_box), ctypes.c_int,
                                       P(ctypes.c_double), ctypes.c_double, ctypes.c_int, P(ctypes.c_int), vp]),
    "cmp_batch_retire": (ctypes.c_int, [vp, ctypes.c_int, vp]),
    "cmp_batch_step": (ctypes.c_int, [vp, vp]),
    "cmp_batch_case_state": (ctypes.c_int, [vp, ctypes.c_int, P(vp), P(ctypes.c_longlong), P(ctypes.c_double),
                                            P(ctypes.c_int), vp]),
    "cmp_batch_case_download": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_double), vp]),
    "cmp_batch_check": (ctypes.c_int, [vp, ctypes.c_int, vp]),
    "cmp_batch_pf_setup": (ctypes.c_int, [vp, P(cmp_pf_params)]),
    "cmp_batch_pf_step": (ctypes.c_int, [vp, ctypes.c_int, dp, vp]),
    "cmp_batch_info": (ctypes.c_int, [vp, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_longlong)]),
    "cmp_level_rhs_noalloc": (ctypes.c_int, [P(cmp_params), P(cmp_geometry), ctypes.c_int, P(cmp_box), ctypes.c_int,
                                             dp, dp, dp, dp, vp]),
}

_lib = None


def load(path=LIB_PATH):
    """Loads libcompas_amd.so and declares its prototypes. Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libcompas_amd.so is not built (%s); run python -m compas_amd.build" % path)
    # torch ships its own HIP runtime with the same soname (libamdhip64.so.7) as /opt/rocm's. Loaded
    # first, it is the one the library binds to; loaded after the library, it would come in as a
    # second runtime in the process and the two would disagree about the device.
    import torch  # noqa: F401
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code, where):
    if code != CMP_OK:
        raise CmpError(code, where)
