"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU work: only host-side calls)."""
import ctypes as C
import glob
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(splash_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    from splash_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 20
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert declared <= set(_lib.SIGNATURES), declared - set(_lib.SIGNATURES)


def test_c_header_is_plain_c():
    """include/splash.h is the boundary any FFI binds: it compiles as strict
    C99 and C11 with every warning an error."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no C compiler")
    header = os.path.join(ROOT, "include", "splash.h")
    for std in ("c99", "c11"):
        r = subprocess.run([cc, "-x", "c", f"-std={std}", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                            header], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_host_mesh_and_geometry_are_bit_identical_to_the_oracle():
    import splash_amd as S
    cases = [("hex", (4,), S.icosphere_hex), ("quad", (4,), S.quad_sphere), ("uv", (32, 16), S.uv_sphere),
             ("ico", (3,), S.icosphere)]
    for kind, args, gen in cases:
        om, pm = O.Mesh(kind, *args), gen(*args)
        assert np.array_equal(om.xyz.view(np.uint64), pm.vertices.view(np.uint64))
        assert np.array_equal(om.foff, pm.face_offsets) and np.array_equal(om.fv, pm.face_vertices)
        og = O.Geometry(om)
        for threads in (4, 0):  # 0: every host thread; the tables must not depend on the count
            pg = S.Geometry(pm, threads=threads)
            for f in ("fc", "normal", "area", "nb_edge", "twin", "ff_plus", "ff_minus", "beta_plus", "beta_minus",
                      "h_plus", "h_minus", "bm_dist", "en"):
                a, b = np.asarray(getattr(og, f)), np.asarray(getattr(pg, f))
                assert a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8)), (kind, f, threads)
            assert og.h0 == pg.h0 and og.round_diff == pg.round_diff


def test_argument_errors_are_codes_not_crashes():
    import splash_amd as S
    from splash_amd._lib import SplashError
    mesh = S.icosphere_hex(1)
    try:
        S.MUSCL_HLLC(mesh, np.ones((mesh.n_faces, 5)), 4, 1.25)
    except ValueError:
        pass
    else:
        raise AssertionError("shape mismatch accepted")
    # dim 4 on an adiabatic solver is rejected before any device work
    try:
        S.MUSCL_HLLC(mesh, np.ones((mesh.n_faces, 4)), 4, 1.25)
    except SplashError as e:
        assert e.code == -1
    else:
        raise AssertionError("dim 4 accepted by the adiabatic solver")
    h = C.c_void_p()
    rc = S._lib.lib().splash_mesh_icosphere_hex(-1, C.byref(h))
    assert rc == -1
    try:
        S.icosphere_hex(99)
    except SplashError as e:
        assert e.code == -1
    else:
        raise AssertionError("level 99 accepted")


def test_gpu_layout_validates_on_every_mesh_family():
    """The stage kernel's patch/halo layout passes its host-side index checks
    (blocks a multiple of 8, own + halo half-edges within the LDS slots,
    twin slots holding the twins) on every generator.  The face order is the
    coordinate bisection: every 64-face leaf fits the 512 slots, so the patch
    count is the minimum, and the halo stays below a fifth of the half-edges
    (runs of the cube-map curve give 16-23% here, the bisection 14-18%)."""
    import splash_amd as S
    for mesh in (S.icosphere_hex(5), S.quad_sphere(5), S.uv_sphere(96, 48), S.icosphere(4), S.icosphere_hex(0)):
        blocks, nonempty, halo, max_halo, max_fset = S.layout_stats(mesh)
        assert blocks % 8 == 0 and 0 < nonempty <= blocks
        assert 0 <= halo < mesh.n_halfedges and max_halo < 512
        assert (halo > 0) == (nonempty > 1)
        assert 0 < max_fset <= 1024
        assert nonempty == -(-mesh.n_faces // 64)
        assert halo < 0.19 * mesh.n_halfedges


def test_gpu_layout_validates_at_every_patch_shape():
    """Patches of up to 128 faces (two face waves) in 128-, 192- and
    256-thread workgroups pass the same host-side index checks on every
    generator; the default cap gives quad and triangle meshes two-wave
    patches in 256-thread workgroups (the 512 slots hold about 110 quads,
    fewer patches and a shorter halo than 64-face patches) and hexagon
    meshes one-wave patches, whose 512 slots fill at about 70 faces."""
    import splash_amd as S
    meshes = (S.icosphere_hex(5), S.quad_sphere(5), S.uv_sphere(96, 48), S.icosphere(4), S.icosphere_hex(0))
    for mesh in meshes:
        for threads in (128, 192, 256):
            for faces in (0, 24, 64, 96, 112, 128):
                blocks, nonempty, halo, max_halo, max_fset = S.layout_stats(mesh, threads, faces)
                assert blocks % 8 == 0 and 0 < nonempty <= blocks
                assert 0 <= halo < mesh.n_halfedges and max_halo < 2 * threads
                assert 0 < max_fset <= 1024
                if faces:
                    assert nonempty >= -(-mesh.n_faces // faces)
    hexm, quad = meshes[0], meshes[1]
    assert S.layout_stats(hexm, 256, 0) == S.layout_stats(hexm, 256, 64)
    assert S.layout_stats(quad, 256, 0) == S.layout_stats(quad, 256, 112)
    assert S.layout_stats(quad, 192, 0) == S.layout_stats(quad, 192, 64)
    one_wave, two_wave = S.layout_stats(quad, 256, 64), S.layout_stats(quad, 256, 112)
    assert two_wave[1] < 0.65 * one_wave[1]   # patches
    assert two_wave[2] < 0.85 * one_wave[2]   # halo half-edges
    for bad in ((64, 64), (256, 129), (256, -1)):
        with pytest.raises(S.SplashError):
            S.layout_stats(quad, *bad)
