"""ctypes bindings to the oracle (oracle/liboracle.so) for the tests.

The oracle is the CPU restatement of the reference solver; tests use it as
the checker for the HIP product path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.environ.get("SPLASH_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")

HLLE, HLLE_P, HLLC, HLLCPLUS = 0, 1, 2, 3


class OrMesh(C.Structure):
    _fields_ = [("nv", C.c_int32), ("nf", C.c_int32), ("xyz", C.POINTER(C.c_double)),
                ("foff", C.POINTER(C.c_int32)), ("fv", C.POINTER(C.c_int32))]


class OrGeom(C.Structure):
    _fields_ = [("nf", C.c_int32), ("nh", C.c_int32), ("nv", C.c_int32),
                ("mesh", C.POINTER(OrMesh)),
                ("fc", C.POINTER(C.c_double)), ("normal", C.POINTER(C.c_double)),
                ("area", C.POINTER(C.c_double)), ("nb_off", C.POINTER(C.c_int32)),
                ("nb", C.POINTER(C.c_int32)), ("nb_edge", C.POINTER(C.c_int32)),
                ("twin", C.POINTER(C.c_int32)), ("en", C.POINTER(C.c_double)),
                ("bm_dist", C.POINTER(C.c_double)), ("ff_plus", C.POINTER(C.c_int32)),
                ("ff_minus", C.POINTER(C.c_int32)), ("beta_plus", C.POINTER(C.c_double)),
                ("beta_minus", C.POINTER(C.c_double)), ("h_plus", C.POINTER(C.c_double)),
                ("h_minus", C.POINTER(C.c_double)), ("h0", C.c_double),
                ("round_diff", C.c_double)]


class OrParams(C.Structure):
    _fields_ = [("accretion_on", C.c_int32), ("friction_on", C.c_int32),
                ("acc_rate", C.c_double), ("e_acc", C.c_double),
                ("omega_acc_abs", C.c_double), ("tilt_angle", C.c_double),
                ("acc_width", C.c_double), ("alpha", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        for name in ("or_uv_sphere",):
            getattr(L, name).restype = C.POINTER(OrMesh)
            getattr(L, name).argtypes = [C.c_int, C.c_int]
        for name in ("or_icosphere", "or_icosphere_hex", "or_quad_sphere"):
            getattr(L, name).restype = C.POINTER(OrMesh)
            getattr(L, name).argtypes = [C.c_int]
        L.or_mesh_free.argtypes = [C.POINTER(OrMesh)]
        L.or_geometry.restype = C.POINTER(OrGeom)
        L.or_geometry.argtypes = [C.POINTER(OrMesh)]
        L.or_geom_free.argtypes = [C.POINTER(OrGeom)]
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(OrMesh), C.POINTER(C.c_double), C.c_int, C.c_int,
                                C.c_double, C.c_double, C.POINTER(OrParams), C.c_int]
        L.or_do_step.argtypes = [C.c_void_p, C.c_double]
        L.or_time.restype = C.c_double
        L.or_time.argtypes = [C.c_void_p]
        L.or_last_dt.restype = C.c_double
        L.or_last_dt.argtypes = [C.c_void_p]
        L.or_area_coeff.restype = C.c_double
        L.or_area_coeff.argtypes = [C.c_void_p]
        L.or_stop_check.restype = C.c_int
        L.or_stop_check.argtypes = [C.c_void_p]
        L.or_get_state.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.or_diagnostics.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.or_curl.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_source.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_reconstruct.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double)]
        L.or_limiter_adiabatic.argtypes = [C.POINTER(C.c_double), C.c_double, C.c_double,
                                           C.c_double, C.c_int, C.POINTER(C.c_double)]
        L.or_flux_star.restype = C.c_int
        L.or_flux_star.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Mesh:
    """Owning wrapper of an oracle mesh, with numpy views."""

    def __init__(self, kind, *args):
        fn = {"uv": lib().or_uv_sphere, "ico": lib().or_icosphere,
              "hex": lib().or_icosphere_hex, "quad": lib().or_quad_sphere}[kind]
        self.ptr = fn(*args)
        if not self.ptr:
            raise ValueError("bad mesh arguments")
        m = self.ptr.contents
        self.nv, self.nf = m.nv, m.nf
        self.xyz = np.ctypeslib.as_array(m.xyz, shape=(m.nv, 3)).copy()
        self.foff = np.ctypeslib.as_array(m.foff, shape=(m.nf + 1,)).copy()
        self.fv = np.ctypeslib.as_array(m.fv, shape=(int(self.foff[-1]),)).copy()

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_mesh_free(self.ptr)
            self.ptr = None

    @property
    def nh(self):
        return int(self.foff[-1])

    def faces(self):
        return [self.fv[self.foff[f]:self.foff[f + 1]] for f in range(self.nf)]


class Geometry:
    """Oracle geometry tables as numpy copies."""

    def __init__(self, mesh):
        p = lib().or_geometry(mesh.ptr)
        g = p.contents
        nf, nh = g.nf, g.nh
        A = np.ctypeslib.as_array
        self.fc = A(g.fc, shape=(nf, 3)).copy()
        self.normal = A(g.normal, shape=(nf, 3)).copy()
        self.area = A(g.area, shape=(nf,)).copy()
        self.nb_off = A(g.nb_off, shape=(nf + 1,)).copy()
        self.nb = A(g.nb, shape=(int(self.nb_off[-1]),)).copy()
        self.nb_edge = A(g.nb_edge, shape=(nh,)).copy()
        self.twin = A(g.twin, shape=(nh,)).copy()
        self.en = A(g.en, shape=(nh, 3)).copy()
        self.bm_dist = A(g.bm_dist, shape=(nh,)).copy()
        self.ff_plus = A(g.ff_plus, shape=(nh, 2)).copy()
        self.ff_minus = A(g.ff_minus, shape=(nh, 2)).copy()
        self.beta_plus = A(g.beta_plus, shape=(nh, 2)).copy()
        self.beta_minus = A(g.beta_minus, shape=(nh, 2)).copy()
        self.h_plus = A(g.h_plus, shape=(nh,)).copy()
        self.h_minus = A(g.h_minus, shape=(nh,)).copy()
        self.h0 = g.h0
        self.round_diff = g.round_diff
        lib().or_geom_free(p)


def make_params(accretion_on=False, friction_on=True, acc_rate=3.25e-7, e_acc=0.02,
                omega_acc_abs=0.4, tilt_angle=0.4, acc_width=0.1, alpha=1e-6):
    """Defaults are the reference's shipped input/parameters.json values."""
    return OrParams(int(accretion_on), int(friction_on), acc_rate, e_acc, omega_acc_abs,
                    tilt_angle, acc_width, alpha)


class Solver:
    """Oracle solver (reference do_step restatement)."""

    def __init__(self, mesh, U, dim, solver, gam, omega_ns=0.0, params=None, threads=1):
        self.mesh = mesh
        self.dim = dim
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (mesh.nf, dim)
        self.params = params if params is not None else make_params()
        self.ptr = lib().or_create(mesh.ptr, _dp(U), dim, solver, gam, omega_ns,
                                   C.byref(self.params), threads)
        if not self.ptr:
            raise ValueError("oracle rejected the solver arguments")

    def do_step(self, dt0):
        lib().or_do_step(self.ptr, dt0)

    def time(self):
        return lib().or_time(self.ptr)

    def last_dt(self):
        return lib().or_last_dt(self.ptr)

    def area_coeff(self):
        return lib().or_area_coeff(self.ptr)

    def get_stop_check(self):
        return bool(lib().or_stop_check(self.ptr))

    def state(self):
        out = np.empty((self.mesh.nf, self.dim))
        lib().or_get_state(self.ptr, _dp(out))
        return out

    def diagnostics(self):
        out = np.empty(5)
        lib().or_diagnostics(self.ptr, _dp(out))
        return out

    def curl(self):
        out = np.empty(self.mesh.nf)
        lib().or_curl(self.ptr, _dp(out))
        return out

    def source(self, fcn, u):
        """source() of p-form state u at unit face centre fcn."""
        fcn = np.ascontiguousarray(fcn, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(5)
        lib().or_source(self.ptr, _dp(fcn), _dp(u), _dp(out))
        return out

    def reconstruct(self, dt):
        """U_plus [nh][dim] of find_U_edges on the current state with step dt."""
        out = np.empty((self.mesh.nh, self.dim))
        lib().or_reconstruct(self.ptr, dt, _dp(out))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_destroy(self.ptr)
            self.ptr = None


def limiter_adiabatic(r, nu, eta_minus, eta_plus, valence):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty(5)
    lib().or_limiter_adiabatic(_dp(r), nu, eta_minus, eta_plus, valence, _dp(out))
    return out


def flux_star(solver, gam, uL, uR, ec, ecn, en):
    uL = np.ascontiguousarray(uL, dtype=np.float64)
    uR = np.ascontiguousarray(uR, dtype=np.float64)
    ec = np.ascontiguousarray(ec, dtype=np.float64)
    en = np.ascontiguousarray(en, dtype=np.float64)
    out = np.zeros(5)
    err = lib().or_flux_star(solver, gam, _dp(uL), _dp(uR), _dp(ec), ecn, _dp(en), _dp(out))
    return out, err
