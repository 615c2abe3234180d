"""Python host mirror of the SPLASH solver classes over the HIP C ABI.

The classes keep the reference's names and constructor arguments
(src/Riemann_solvers/HLLE.hpp:13, HLLE_p.hpp:11, HLLC.hpp:11,
HLLCplus.hpp:11) and its methods do_step / time / get_stop_check
(src/MUSCL_base/MUSCL_base.hpp:97, :227, :232); every call goes through
libsplash_amd.so.  Extensions: get_state, diagnostics, do_steps, last_dt.
"""
import ctypes as C

import numpy as np

from . import ic
from ._lib import (FP64, HLLC, HLLCPLUS, HLLE, HLLE_P, MIXED, GeometryView, MeshDesc, Params,
                   SplashError, check, lib)

__all__ = ["SurfaceMesh", "uv_sphere", "icosphere", "icosphere_hex", "quad_sphere", "Geometry",
           "MUSCL_HLLE", "MUSCL_HLLE_p", "MUSCL_HLLC", "MUSCL_HLLCplus", "Ensemble", "default_params",
           "SplashError", "FP64", "MIXED", "ic"]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class SurfaceMesh:
    """Flattened polygon mesh (vertices + per-face vertex lists in pmp order)."""

    def __init__(self, vertices, face_offsets, face_vertices):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        self.face_offsets = np.ascontiguousarray(face_offsets, dtype=np.int32)
        self.face_vertices = np.ascontiguousarray(face_vertices, dtype=np.int32)

    @property
    def n_faces(self):
        return len(self.face_offsets) - 1

    @property
    def n_vertices(self):
        return len(self.vertices)

    @property
    def n_halfedges(self):
        return int(self.face_offsets[-1])

    def desc(self):
        return MeshDesc(self.n_vertices, self.n_faces, _dp(self.vertices), _ip(self.face_offsets),
                        _ip(self.face_vertices))

    def face_centres(self):
        """Vertex-mean face centres (MUSCL_geometry.hpp:176-187)."""
        fo = self.face_offsets
        val = np.diff(fo)
        sums = np.add.reduceat(self.vertices[self.face_vertices], fo[:-1], axis=0)
        return sums / val[:, None]


def _generated(fn, *args):
    h = C.c_void_p()
    check(fn(*args, C.byref(h)))
    try:
        d = MeshDesc()
        check(lib().splash_mesh_view(h, C.byref(d)))
        nv, nf = d.n_vertices, d.n_faces
        verts = np.ctypeslib.as_array(d.vertices, shape=(nv, 3)).copy()
        fo = np.ctypeslib.as_array(d.face_offsets, shape=(nf + 1,)).copy()
        fv = np.ctypeslib.as_array(d.face_vertices, shape=(int(fo[-1]),)).copy()
    finally:
        lib().splash_mesh_free(h)
    return SurfaceMesh(verts, fo, fv)


def uv_sphere(n_slices, n_stacks):
    """sph_gen.h:215-265."""
    return _generated(lib().splash_mesh_uv_sphere, n_slices, n_stacks)


def icosphere(n):
    """sph_gen.h:267-276."""
    return _generated(lib().splash_mesh_icosphere, n)


def icosphere_hex(n):
    """sph_gen.h:278-284."""
    return _generated(lib().splash_mesh_icosphere_hex, n)


def quad_sphere(n):
    """sph_gen.h:286-295."""
    return _generated(lib().splash_mesh_quad_sphere, n)


class Geometry:
    """Host stencil tables of the product (copies), for checks."""

    def __init__(self, mesh, threads=0):
        h = C.c_void_p()
        d = mesh.desc()
        check(lib().splash_geometry_create(C.byref(d), threads, C.byref(h)))
        try:
            v = GeometryView()
            check(lib().splash_geometry_get_view(h, C.byref(v)))
            nf, nh = v.nf, v.nh
            A = np.ctypeslib.as_array
            self.fc = A(v.fc, shape=(nf, 3)).copy()
            self.normal = A(v.normal, shape=(nf, 3)).copy()
            self.area = A(v.area, shape=(nf,)).copy()
            self.nb_edge = A(v.nb_edge, shape=(nh,)).copy()
            self.twin = A(v.twin, shape=(nh,)).copy()
            self.ff_plus = A(v.ff_plus, shape=(nh, 2)).copy()
            self.ff_minus = A(v.ff_minus, shape=(nh, 2)).copy()
            self.beta_plus = A(v.beta_plus, shape=(nh, 2)).copy()
            self.beta_minus = A(v.beta_minus, shape=(nh, 2)).copy()
            self.h_plus = A(v.h_plus, shape=(nh,)).copy()
            self.h_minus = A(v.h_minus, shape=(nh,)).copy()
            self.bm_dist = A(v.bm_dist, shape=(nh,)).copy()
            self.en = A(v.en, shape=(nh, 3)).copy()
            self.h0 = v.h0
            self.round_diff = v.round_diff
        finally:
            lib().splash_geometry_free(h)


def layout_stats(mesh, block_threads=None, patch_faces=None):
    """(blocks, non-empty blocks, halo half-edges, max halo, max LDS face set)
    of the validated internal GPU layout of `mesh` (host only): 64-face
    patches in 256-thread workgroups, or the given shape (patch_faces 0: the
    mesh's default cap)."""
    out = (C.c_int32 * 5)()
    d = mesh.desc()
    if block_threads is None and patch_faces is None:
        check(lib().splash_layout_stats(C.byref(d), out))
    else:
        check(lib().splash_layout_stats_shape(C.byref(d), int(block_threads or 0), int(patch_faces or 0), out))
    return tuple(out)


def default_params(**overrides):
    """The reference's shipped input/parameters.json physics keys, accretion off."""
    p = Params()
    lib().splash_default_params(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class _MUSCL:
    _solver = None
    _dim = None

    def __init__(self, mesh, U_in, dim, gam, omega_ns=0.0, accretion_on=False, threads=1,
                 params=None, precision=FP64):
        U = np.ascontiguousarray(U_in, dtype=np.float64)
        if U.ndim != 2 or U.shape[0] != mesh.n_faces or U.shape[1] != dim:
            raise ValueError(f"U_in must be [{mesh.n_faces}][{dim}], got {U.shape}")
        p = params if params is not None else default_params()
        p.accretion_on = int(bool(accretion_on))
        p.precision = int(precision)
        self._mesh = mesh
        self._dim = dim
        self._ctx = C.c_void_p()
        d = mesh.desc()
        check(lib().splash_create(C.byref(d), _dp(U), dim, self._solver, gam, omega_ns, C.byref(p),
                                  C.byref(self._ctx)))

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            lib().splash_destroy(ctx)
            self._ctx = None

    def do_step(self, dt0):
        check(lib().splash_do_step(self._ctx, dt0))

    def do_steps(self, dt0, n):
        check(lib().splash_do_steps(self._ctx, dt0, int(n)))

    def profile_step(self, dt0, n):
        """Mean ms per launch of [dt finalize, stage 1, stage 2, whole step]
        over n steps launched back to back on the run's stream, each kernel
        with start/stop HIP events bound to its dispatch (MIXED launches no
        dt kernel: its stage 1 forms dt, and the first entry is 0)."""
        out = np.zeros(4)
        check(lib().splash_profile_step(self._ctx, dt0, int(n), _dp(out)))
        return out

    def synchronize(self):
        check(lib().splash_synchronize(self._ctx))

    def time(self):
        t = C.c_double()
        check(lib().splash_time(self._ctx, C.byref(t)))
        return t.value

    def last_dt(self):
        t = C.c_double()
        check(lib().splash_last_dt(self._ctx, C.byref(t)))
        return t.value

    def get_stop_check(self):
        s = C.c_int()
        check(lib().splash_stop_check(self._ctx, C.byref(s)))
        return bool(s.value)

    def stop_bits(self):
        b = C.c_int()
        first = C.c_uint64()
        check(lib().splash_stop_bits(self._ctx, C.byref(b), C.byref(first)))
        return b.value, first.value

    def get_state(self):
        out = np.empty((self._mesh.n_faces, self._dim))
        check(lib().splash_get_state(self._ctx, _dp(out)))
        return out

    def diagnostics(self):
        out = np.empty(5)
        check(lib().splash_diagnostics(self._ctx, _dp(out)))
        return out

    def set_state(self, U, t, stop_bits=0):
        """Resume from a checkpoint: state (mesh face order), time, stop bits."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != (self._mesh.n_faces, self._dim):
            raise ValueError(f"U must be [{self._mesh.n_faces}][{self._dim}], got {U.shape}")
        check(lib().splash_set_state(self._ctx, _dp(U), float(t), int(stop_bits)))

    def get_light_curves(self):
        """adiabatic.hpp:228-333: [flux_0, flux_45, flux_90, flux_180]."""
        out = np.empty(4)
        check(lib().splash_light_curves(self._ctx, _dp(out)))
        return out

    def curl(self):
        """write_t_curl's per-face values (adiabatic.hpp:135-178,
        isothermal.hpp:72-110) of the current state, [n_faces]."""
        out = np.empty(self._mesh.n_faces)
        check(lib().splash_curl(self._ctx, _dp(out)))
        return out

    def patch_faces(self):
        """Faces per stage-kernel patch (mixed precision; 0 for FP64)."""
        n = C.c_int32()
        check(lib().splash_patch_faces(self._ctx, C.byref(n)))
        return n.value

    def patch_shape(self):
        """(faces per patch, threads per workgroup) of the stage kernel
        (mixed precision; (0, 0) for FP64)."""
        f, t = C.c_int32(), C.c_int32()
        check(lib().splash_patch_shape(self._ctx, C.byref(f), C.byref(t)))
        return f.value, t.value


class MUSCL_HLLE(_MUSCL):
    """MUSCL_HLLE(mesh, U_in, dim, gam, threads) -- HLLE.hpp:13, isothermal dim 4."""
    _solver = HLLE

    def __init__(self, mesh, U_in, dim, gam, threads=1, params=None, precision=FP64):
        super().__init__(mesh, U_in, dim, gam, 0.0, False, threads, params, precision)


class MUSCL_HLLE_p(_MUSCL):
    """MUSCL_HLLE_p(mesh, U_in, dim, gam, omega_ns, accretion_on, threads) -- HLLE_p.hpp:11."""
    _solver = HLLE_P


class MUSCL_HLLC(_MUSCL):
    """MUSCL_HLLC(mesh, U_in, dim, gam, omega_ns, accretion_on, threads) -- HLLC.hpp:11."""
    _solver = HLLC


class MUSCL_HLLCplus(_MUSCL):
    """MUSCL_HLLCplus(mesh, U_in, dim, gam, omega_ns, accretion_on, threads) -- HLLCplus.hpp:11."""
    _solver = HLLCPLUS


class Ensemble:
    """Independent runs of one solver advanced together (splash_ensemble_*).

    runs: list of (mesh, U_in, gam, omega_ns, params); every RK stage is one
    launch of the FP32-flux / FP64-state kernels for all runs, and each run
    evolves as MUSCL_*(..., precision=MIXED) would on the same inputs."""

    def __init__(self, runs, dim, solver):
        n = len(runs)
        self._descs = (MeshDesc * n)(*[m.desc() for m, *_ in runs])
        self._meshes = [m for m, *_ in runs]
        self._U = [np.ascontiguousarray(U, dtype=np.float64) for _, U, *_ in runs]
        for (m, *_), U in zip(runs, self._U):
            if U.shape != (m.n_faces, dim):
                raise ValueError(f"U_in must be [{m.n_faces}][{dim}], got {U.shape}")
        ptrs = (C.POINTER(C.c_double) * n)(*[_dp(U) for U in self._U])
        gam = np.array([r[2] for r in runs], dtype=np.float64)
        om = np.array([r[3] for r in runs], dtype=np.float64)
        params = (Params * n)(*[r[4] if r[4] is not None else default_params() for r in runs])
        self._dim = dim
        self._e = C.c_void_p()
        check(lib().splash_ensemble_create(n, self._descs, ptrs, dim, solver, _dp(gam), _dp(om), params,
                                           C.byref(self._e)))
        self.n_runs = n

    def __del__(self):
        e = getattr(self, "_e", None)
        if e:
            lib().splash_ensemble_destroy(e)
            self._e = None

    def _dt0(self, dt0):
        a = np.broadcast_to(np.asarray(dt0, dtype=np.float64), (self.n_runs,))
        return np.ascontiguousarray(a)

    def do_steps(self, dt0, n):
        """n steps of every run; dt0 is a scalar or one value per run."""
        d = self._dt0(dt0)
        check(lib().splash_ensemble_step(self._e, _dp(d), int(n)))

    def profile_step(self, dt0, n):
        d = self._dt0(dt0)
        out = np.zeros(4)
        check(lib().splash_ensemble_profile_step(self._e, _dp(d), int(n), _dp(out)))
        return out

    def synchronize(self):
        check(lib().splash_ensemble_synchronize(self._e))

    def info(self):
        """(n_runs, total faces, total half-edges, stage-kernel workgroups)."""
        r, f, h, b = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().splash_ensemble_info(self._e, C.byref(r), C.byref(f), C.byref(h), C.byref(b)))
        return r.value, f.value, h.value, b.value

    def patch_faces(self):
        n = C.c_int32()
        check(lib().splash_ensemble_patch_faces(self._e, C.byref(n)))
        return n.value

    def patch_shape(self):
        f, t = C.c_int32(), C.c_int32()
        check(lib().splash_ensemble_patch_shape(self._e, C.byref(f), C.byref(t)))
        return f.value, t.value

    def time(self, run):
        t = C.c_double()
        check(lib().splash_ensemble_time(self._e, int(run), C.byref(t)))
        return t.value

    def last_dt(self, run):
        t = C.c_double()
        check(lib().splash_ensemble_last_dt(self._e, int(run), C.byref(t)))
        return t.value

    def get_stop_check(self, run):
        s = C.c_int()
        check(lib().splash_ensemble_stop_check(self._e, int(run), C.byref(s)))
        return bool(s.value)

    def set_state(self, run, U, t, stop_bits=0):
        """Resume run `run`: state (mesh face order), time, stop bits."""
        n = C.c_int32()
        check(lib().splash_ensemble_num_faces(self._e, int(run), C.byref(n)))
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != (n.value, self._dim):
            raise ValueError(f"state must be [{n.value}][{self._dim}]")
        check(lib().splash_ensemble_set_state(self._e, int(run), _dp(U), float(t), int(stop_bits)))

    def diagnostics(self, run):
        """Run `run`'s diagnostics sums (one device readout of every run)."""
        out = np.empty(5)
        check(lib().splash_ensemble_diagnostics(self._e, int(run), _dp(out)))
        return out

    def get_light_curves(self, run):
        out = np.empty(4)
        check(lib().splash_ensemble_light_curves(self._e, int(run), _dp(out)))
        return out

    def get_state(self, run):
        n = C.c_int32()
        check(lib().splash_ensemble_num_faces(self._e, int(run), C.byref(n)))
        out = np.empty((n.value, self._dim))
        check(lib().splash_ensemble_get_state(self._e, int(run), _dp(out)))
        return out
