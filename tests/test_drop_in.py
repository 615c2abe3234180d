"""The C++ drop-in classes (include/splash/MUSCL.hpp) driven by the
reference's own driver loop (tests/cpp/drop_in_driver.cpp restating
test.cpp:9-152): parameter-file behaviour on the CPU, the full loop on the
GPU against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from lightcurve_ref import light_curves

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "drop_in_driver")
GAM3D = 4 / 3
GAM = 2 - 1 / GAM3D

# the reference's shipped input/parameters.json keys (driver + physics)
PARAMS = {"dim": 5, "dt": 1e-4, "t_max": 6000, "maxstep": 20, "skipstep": 5, "gam3d": GAM3D, "omega_ns": 0.0,
          "threads": 8, "accretion_on": False, "accretion_rate": 3.25e-7, "e_acc": 0.02, "omega_acc_abs": 0.4,
          "tilt_angle": 0.4, "acc_width": 0.1, "friction_on": True, "alpha": 1e-6}


def _setup(tmp_path, params, U=None):
    (tmp_path / "input").mkdir()
    (tmp_path / "results").mkdir()
    (tmp_path / "input" / "parameters.json").write_text(json.dumps(params, indent=1))
    if U is not None:
        np.savetxt(tmp_path / "input" / "input.dat", U, fmt="%.17g")


def test_missing_physics_keys_raise_before_any_device_work(tmp_path):
    """adiabatic.hpp:59-68 reads the physics keys and throws when they are
    absent; the drop-in does the same before touching the GPU."""
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    p = {k: v for k, v in PARAMS.items() if k != "accretion_rate"}
    _setup(tmp_path, p)
    r = subprocess.run([DRIVER, "quad", "2"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "parse error" in (r.stderr + r.stdout)


@pytest.mark.gpu
def test_driver_loop_matches_oracle(tmp_path):
    import splash_amd as S
    mesh = S.quad_sphere(4)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=11, omega=(0, 0, 0.2))
    _setup(tmp_path, PARAMS, U0)
    r = subprocess.run([DRIVER, "quad", "4"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("t= ")]
    assert len(lines) == PARAMS["maxstep"]
    # text outputs carry 6 significant digits (std::ostream default)
    final = np.loadtxt(tmp_path / "results" / "final_state.dat")
    ref = O.Solver(O.Mesh("quad", 4), np.loadtxt(tmp_path / "input" / "input.dat"), 5, O.HLLCPLUS, GAM, 0.0)
    curl_rows = []
    for k in range(PARAMS["maxstep"]):
        ref.do_step(PARAMS["dt"])
        if k % PARAMS["skipstep"] == 0:
            curl_rows.append(np.concatenate([[ref.time()], ref.curl()]))
    assert np.allclose(final, ref.state(), rtol=1e-5, atol=1e-6 * np.max(np.abs(ref.state())))
    rho = np.loadtxt(tmp_path / "results" / "rho.dat")
    assert rho.shape == (1 + PARAMS["maxstep"] // PARAMS["skipstep"], 1 + mesh.n_faces)
    # write_t_curl after every skipstep-th step, as test.cpp:123-131
    curl = np.loadtxt(tmp_path / "results" / "curl.dat")
    want = np.array(curl_rows)
    assert curl.shape == want.shape
    assert np.allclose(curl, want, rtol=1e-5, atol=1e-5 * np.max(np.abs(want[:, 1:])))
    fc = np.loadtxt(tmp_path / "results" / "face_centers.dat")
    assert fc.shape == (mesh.n_faces, 3) and np.allclose(np.linalg.norm(fc, axis=1), 1, atol=1e-14)
    # the four light curves written every step (get_light_curves on the
    # device), as test.cpp:110-122
    g = S.Geometry(mesh)
    want = light_curves(ref.state(), g.fc, g.area, GAM)
    for k, name in enumerate(("lightcurve0", "lightcurve45", "lightcurve90", "lightcurve180")):
        lc = np.loadtxt(tmp_path / "results" / f"{name}.dat")
        assert lc.shape == (1 + PARAMS["maxstep"], 2)
        assert lc[-1, 0] == pytest.approx(ref.time(), rel=1e-5) and lc[-1, 1] == pytest.approx(want[k], rel=1e-5)
    # the writers' other series against values recomputed from the oracle's
    # states at the same steps (initial state, then after steps 1, 6, 11, 16)
    ref2 = O.Solver(O.Mesh("quad", 4), np.loadtxt(tmp_path / "input" / "input.dat"), 5, O.HLLCPLUS, GAM, 0.0)
    states, times = [ref2.state()], [0.0]
    for k in range(PARAMS["maxstep"]):
        ref2.do_step(PARAMS["dt"])
        if k % PARAMS["skipstep"] == 0:
            states.append(ref2.state())
            times.append(ref2.time())
    fc = g.fc
    r = fc / np.linalg.norm(fc, axis=1)[:, None]
    theta = np.arccos(fc[:, 2] / np.linalg.norm(fc, axis=1))

    def rows(fn):
        return np.array([np.concatenate([[t], fn(U)]) for t, U in zip(times, states)])

    def vel(U):
        return np.cross(r, U[:, 1:4]) / (-U[:, :1])

    def pressure(U):  # adiabatic.hpp:113-133 with :582-592
        v = vel(U)
        return np.maximum(1e-16, (U[:, 4] - U[:, 0] * np.sum(v * v, axis=1) / 2) * (GAM - 1))

    def omega_z(U):  # adiabatic.hpp:195-226
        with np.errstate(divide="ignore", invalid="ignore"):
            om = np.cross(r, vel(U))[:, 2] / (np.sin(theta) * np.sin(theta))
        om[~np.isfinite(om) | (np.abs(om) > 1e4)] = 0
        return om

    series = {"p": pressure, "omega": omega_z, "Lx": lambda U: U[:, 1], "Ly": lambda U: U[:, 2],
              "Lz": lambda U: U[:, 3], "rho": lambda U: U[:, 0]}
    for name, fn in series.items():
        got = np.loadtxt(tmp_path / "results" / f"{name}.dat")
        want = rows(fn)
        assert got.shape == want.shape, name
        scale = np.max(np.abs(want[:, 1:]))
        assert np.allclose(got[:, 1:], want[:, 1:], rtol=1e-5, atol=1e-5 * scale), name
        assert np.allclose(got[:, 0], want[:, 0], rtol=1e-5, atol=1e-12), name
    # the mesh writers (MUSCL_geometry.hpp:1069-1117)
    faces = [ln.split() for ln in (tmp_path / "results" / "faces.dat").read_text().splitlines()]
    fo, fv = mesh.face_offsets, mesh.face_vertices
    assert [[int(x) for x in f] for f in faces] == [list(map(int, fv[fo[f]:fo[f + 1]])) for f in range(mesh.n_faces)]
    verts = np.loadtxt(tmp_path / "results" / "vertices.dat")
    assert np.allclose(verts, mesh.vertices, rtol=0, atol=1e-14)


def test_pmp_mesh_flattening():
    """from_pmp() (include/splash/MUSCL.hpp) turns an object with pmp's
    SurfaceMesh interface into exactly the flat arrays of the same mesh."""
    prog = os.path.join(ROOT, "tests", "cpp", "pmp_adapter")
    r = subprocess.run([prog, "flatten"], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_pmp_mesh_constructors(tmp_path):
    """The drop-in constructors called with a pmp-style mesh step exactly as
    when called with the flattened mesh (HLLC+ from input/parameters.json,
    and HLLE)."""
    (tmp_path / "input").mkdir()
    (tmp_path / "results").mkdir()
    (tmp_path / "input" / "parameters.json").write_text(json.dumps(PARAMS, indent=1))
    prog = os.path.join(ROOT, "tests", "cpp", "pmp_adapter")
    r = subprocess.run([prog, "solve"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_checkpoint_round_trip(tmp_path, precision):
    """write_checkpoint after 13 steps, 11 more steps; a second solver that
    reads the checkpoint and steps 11 ends bit for bit in the same state."""
    (tmp_path / "results").mkdir()
    prog = os.path.join(ROOT, "tests", "cpp", "checkpoint_roundtrip")
    r = subprocess.run([prog, str(precision), str(tmp_path / "ck.bin")], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_ensemble_wrapper_matches_single_solvers(tmp_path):
    (tmp_path / "results").mkdir()
    prog = os.path.join(ROOT, "tests", "cpp", "ensemble_vs_single")
    r = subprocess.run([prog], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


GOLDEN_PARAMS = os.path.join(os.path.dirname(__file__), "golden", "parameters.json")
PHYSICS_KEYS = {"accretion_rate": "acc_rate", "e_acc": "e_acc", "omega_acc_abs": "omega_acc_abs",
                "tilt_angle": "tilt_angle", "acc_width": "acc_width", "alpha": "alpha",
                "friction_on": "friction_on"}


def test_reference_parameter_file_is_the_default_physics():
    """tests/golden/parameters.json is the reference's shipped input file.
    The drop-in reader parses it to exactly its values (the keys of
    adiabatic.hpp:59-68), and those are the library's and the oracle's
    default physics parameters."""
    import splash_amd as S
    with open(GOLDEN_PARAMS) as f:
        ref = json.load(f)
    reader = os.path.join(os.path.dirname(__file__), "cpp", "params_reader")
    r = subprocess.run([reader, GOLDEN_PARAMS], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    parsed = {k: float(v) for k, v in (ln.split() for ln in r.stdout.splitlines())}
    d = S.default_params()
    o = O.make_params()
    for key, field in PHYSICS_KEYS.items():
        want = float(ref[key])
        assert parsed[key] == want, key
        assert float(getattr(d, field)) == want, key
        assert float(getattr(o, field)) == want, key
