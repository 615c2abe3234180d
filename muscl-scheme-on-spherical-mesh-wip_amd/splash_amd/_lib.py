"""ctypes loader for libsplash_amd.so (the C ABI in include/splash.h).

The library is built in-tree by the package Makefile; there is no fallback:
importing the solver classes without the built library raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SPLASH_AMD_LIB selects another build of the same library (for example the
# sanitizer build of the host code); the default is the in-tree build
LIB_PATH = os.environ.get("SPLASH_AMD_LIB") or os.path.join(HERE, "libsplash_amd.so")

SPLASH_OK = 0
ERRORS = {-1: "EINVAL", -2: "EMESH", -3: "EDEVICE", -4: "ENOMEM", -5: "EINTERNAL"}

HLLE, HLLE_P, HLLC, HLLCPLUS = 0, 1, 2, 3
FP64, MIXED = 0, 1


class SplashError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"splash error {ERRORS.get(code, code)}: {message}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_int32), ("n_faces", C.c_int32),
                ("vertices", C.POINTER(C.c_double)), ("face_offsets", C.POINTER(C.c_int32)),
                ("face_vertices", C.POINTER(C.c_int32))]


class Params(C.Structure):
    _fields_ = [("accretion_on", C.c_int32), ("friction_on", C.c_int32),
                ("acc_rate", C.c_double), ("e_acc", C.c_double), ("omega_acc_abs", C.c_double),
                ("tilt_angle", C.c_double), ("acc_width", C.c_double), ("alpha", C.c_double),
                ("precision", C.c_int32), ("setup_threads", C.c_int32), ("patch_faces", C.c_int32),
                ("no_graph", C.c_int32), ("block_threads", C.c_int32)]


class GeometryView(C.Structure):
    _fields_ = [("nf", C.c_int32), ("nh", C.c_int32),
                ("fc", C.POINTER(C.c_double)), ("normal", C.POINTER(C.c_double)),
                ("area", C.POINTER(C.c_double)), ("nb_edge", C.POINTER(C.c_int32)),
                ("twin", C.POINTER(C.c_int32)), ("ff_plus", C.POINTER(C.c_int32)),
                ("ff_minus", C.POINTER(C.c_int32)), ("beta_plus", C.POINTER(C.c_double)),
                ("beta_minus", C.POINTER(C.c_double)), ("h_plus", C.POINTER(C.c_double)),
                ("h_minus", C.POINTER(C.c_double)), ("bm_dist", C.POINTER(C.c_double)),
                ("en", C.POINTER(C.c_double)), ("h0", C.c_double), ("round_diff", C.c_double)]


# every exported entry point of include/splash.h with its ctypes signature
SIGNATURES = {
    "splash_default_params": (None, [C.POINTER(Params)]),
    "splash_mesh_uv_sphere": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "splash_mesh_icosphere": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "splash_mesh_icosphere_hex": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "splash_mesh_quad_sphere": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "splash_mesh_view": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc)]),
    "splash_mesh_free": (None, [C.c_void_p]),
    "splash_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_double), C.c_int, C.c_int,
                                C.c_double, C.c_double, C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "splash_destroy": (None, [C.c_void_p]),
    "splash_do_step": (C.c_int, [C.c_void_p, C.c_double]),
    "splash_do_steps": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "splash_profile_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double)]),
    "splash_synchronize": (C.c_int, [C.c_void_p]),
    "splash_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "splash_last_dt": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "splash_stop_check": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "splash_stop_bits": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "splash_get_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "splash_set_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_double, C.c_int]),
    "splash_diagnostics": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "splash_light_curves": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "splash_curl": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "splash_num_faces": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "splash_face_geometry": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "splash_geometry_create": (C.c_int, [C.POINTER(MeshDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "splash_geometry_get_view": (C.c_int, [C.c_void_p, C.POINTER(GeometryView)]),
    "splash_geometry_free": (None, [C.c_void_p]),
    "splash_layout_stats": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_int32)]),
    "splash_layout_stats_shape": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "splash_ensemble_create": (C.c_int, [C.c_int, C.POINTER(MeshDesc), C.POINTER(C.POINTER(C.c_double)), C.c_int,
                                         C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(Params),
                                         C.POINTER(C.c_void_p)]),
    "splash_ensemble_destroy": (None, [C.c_void_p]),
    "splash_ensemble_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int]),
    "splash_ensemble_profile_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int,
                                               C.POINTER(C.c_double)]),
    "splash_ensemble_synchronize": (C.c_int, [C.c_void_p]),
    "splash_ensemble_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "splash_ensemble_num_faces": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32)]),
    "splash_ensemble_patch_faces": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "splash_patch_faces": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "splash_patch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "splash_ensemble_patch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "splash_ensemble_time": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "splash_ensemble_last_dt": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "splash_ensemble_stop_check": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "splash_ensemble_get_state": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "splash_ensemble_set_state": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_double, C.c_int]),
    "splash_ensemble_diagnostics": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "splash_ensemble_light_curves": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "splash_readout_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "splash_copy_bandwidth": (C.c_int, [C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "splash_last_error_message": (C.c_char_p, []),
}

_lib = None


def lib():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C <package>/csrc` or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != SPLASH_OK:
        msg = lib().splash_last_error_message()
        raise SplashError(code, msg.decode() if msg else "")
    return code
