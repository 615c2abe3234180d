"""Batched ensemble (splash_ensemble_*) against single runs.

Every run of an ensemble is advanced by the same kernels as a single
MIXED-precision solver on the same inputs, with its own dt clamp, time and
stop_check; the bar is bit-for-bit equality with the single runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAM = 2 - 1 / (4 / 3)


def _runs(S):
    acc = S.default_params(accretion_on=1, friction_on=1, acc_rate=1e-3, e_acc=0.1, omega_acc_abs=0.8,
                           tilt_angle=0.3, acc_width=0.1, alpha=1e-6)
    spec = [(S.icosphere_hex(3), 0.0, None, (0, 0, 0.3), 1e-4),
            (S.icosphere_hex(4), 0.1, acc, (0.2, 0, 0.1), 2e-4),
            (S.quad_sphere(4), 0.0, None, (0, 0.3, 0), 1e-4),
            (S.uv_sphere(32, 16), 0.2, acc, (0.1, 0.1, 0.1), 5e-5),
            (S.icosphere(3), 0.0, None, (0, 0, 0.5), 1e-4)]
    out = []
    for k, (m, om, p, omega, dt0) in enumerate(spec):
        U0 = S.ic.adiabatic_state(m.face_centres(), GAM - 0.01 * k, seed=10 + k, omega=omega)
        out.append((m, U0, GAM - 0.01 * k, om, p, dt0))
    return out


def test_ensemble_runs_equal_single_mixed_runs():
    import splash_amd as S
    runs = _runs(S)
    ens = S.Ensemble([(m, U0, g, om, p) for m, U0, g, om, p, _ in runs], 5, S.HLLCPLUS)
    n_runs, faces, halfedges, blocks = ens.info()
    assert n_runs == len(runs) and faces == sum(m.n_faces for m, *_ in runs) and blocks % 8 == 0
    dt0 = [r[5] for r in runs]
    singles = []
    for m, U0, g, om, p, d in runs:
        s = S.MUSCL_HLLCplus(m, U0, 5, g, om, p is not None, params=p, precision=S.MIXED)
        singles.append(s)
    for n in (19, 8, 3):  # graph replays plus eager remainders
        ens.do_steps(dt0, n)
        for s, d in zip(singles, dt0):
            s.do_steps(d, n)
    for r, (s, d) in enumerate(zip(singles, dt0)):
        A, B = s.get_state(), ens.get_state(r)
        print(f"run {r}: max |diff| {np.max(np.abs(A - B)):.3e}, t {ens.time(r):.6g}")
        assert np.array_equal(A, B)
        assert ens.time(r) == s.time()
        assert ens.last_dt(r) == s.last_dt()
        assert ens.get_stop_check(r) == s.get_stop_check()


def test_ensemble_readout_equals_single_runs():
    """Each run's diagnostics sums and light curves from the ensemble's one
    readout (a run's faces in blocks from its own first face, its own gam)
    equal a single solver's bit for bit, before stepping and after; the
    readout follows the steps."""
    import splash_amd as S
    runs = _runs(S)
    ens = S.Ensemble([(m, U0, g, om, p) for m, U0, g, om, p, _ in runs], 5, S.HLLCPLUS)
    singles = [S.MUSCL_HLLCplus(m, U0, 5, g, om, p is not None, params=p, precision=S.MIXED)
               for m, U0, g, om, p, _ in runs]
    dt0 = [r[5] for r in runs]
    before = [ens.get_light_curves(r) for r in range(len(runs))]
    for step in (0, 11):
        ens.do_steps(dt0, step)
        for s, d in zip(singles, dt0):
            s.do_steps(d, step)
        for r, s in enumerate(singles):
            assert np.array_equal(ens.diagnostics(r), s.diagnostics())
            assert np.array_equal(ens.get_light_curves(r), s.get_light_curves())
            assert ens.time(r) == s.time()
    assert not any(np.array_equal(ens.get_light_curves(r), before[r]) for r in range(len(runs)))


def test_ensemble_set_state_resumes_exactly():
    """A fresh ensemble whose runs are resumed one by one from another's
    states, times and stop_check after 13 steps continues bit for bit as the
    uninterrupted one over 11 more steps (graph replay plus eager steps)."""
    import splash_amd as S
    runs = _runs(S)
    spec = [(m, U0, g, om, p) for m, U0, g, om, p, _ in runs]
    dt0 = [r[5] for r in runs]
    a = S.Ensemble(spec, 5, S.HLLCPLUS)
    a.do_steps(dt0, 13)
    saved = [(a.get_state(r), a.time(r), int(a.get_stop_check(r))) for r in range(len(runs))]
    b = S.Ensemble(spec, 5, S.HLLCPLUS)
    for r, (U, t, stop) in enumerate(saved):
        b.set_state(r, U, t, stop)
    a.do_steps(dt0, 11)
    b.do_steps(dt0, 11)
    for r in range(len(runs)):
        assert np.array_equal(a.get_state(r), b.get_state(r))
        assert a.time(r) == b.time(r) and a.last_dt(r) == b.last_dt(r)
        assert a.get_stop_check(r) == b.get_stop_check(r)


def test_ensemble_of_many_small_runs():
    """120 runs of 42 to 642 faces (icosphere_hex levels 1-3, a few patches
    each, many sharing one 8-block padding group): every run's state, time
    and light curves equal those of its single solver bit for bit."""
    import splash_amd as S
    meshes = [S.icosphere_hex(1 + k % 3) for k in range(3)]
    runs = []
    for k in range(120):
        m = meshes[k % 3]
        U0 = S.ic.adiabatic_state(m.face_centres(), GAM, seed=100 + k, omega=(0.0, 0.1 * (k % 5), 0.3))
        runs.append((m, U0, GAM - 0.001 * (k % 7), 0.0, None))
    ens = S.Ensemble(runs, 5, S.HLLC)
    dt0 = [1e-4 * (1 + k % 4) for k in range(len(runs))]
    ens.do_steps(dt0, 21)
    for r in range(0, len(runs), 7):
        m, U0, g, om, p = runs[r]
        s = S.MUSCL_HLLC(m, U0, 5, g, om, False, params=p, precision=S.MIXED)
        s.do_steps(dt0[r], 21)
        assert np.array_equal(ens.get_state(r), s.get_state())
        assert ens.time(r) == s.time()
        assert np.array_equal(ens.get_light_curves(r), s.get_light_curves())


def test_ensemble_isothermal_and_changing_dt0():
    """dim 4 / MUSCL_HLLE runs, and a per-run dt0 that changes between calls."""
    import splash_amd as S
    meshes = [S.icosphere_hex(3), S.quad_sphere(3)]
    U = [S.ic.isothermal_state(m.face_centres(), seed=3 + k, sigma_amp=0.05) for k, m in enumerate(meshes)]
    ens = S.Ensemble([(m, u, GAM, 0.0, None) for m, u in zip(meshes, U)], 4, S.HLLE)
    singles = [S.MUSCL_HLLE(m, u, 4, GAM, precision=S.MIXED) for m, u in zip(meshes, U)]
    for dt0 in ([1e-3, 5e-4], [2e-4, 2e-4]):
        ens.do_steps(dt0, 12)
        for s, d in zip(singles, dt0):
            s.do_steps(d, 12)
    for r, s in enumerate(singles):
        assert np.array_equal(s.get_state(), ens.get_state(r))
        assert ens.time(r) == s.time()


def test_ensemble_rejects_mixed_solver_shapes():
    import splash_amd as S
    m = S.icosphere_hex(2)
    U4 = S.ic.isothermal_state(m.face_centres(), seed=1)
    with pytest.raises(S.SplashError):
        S.Ensemble([(m, U4, GAM, 0.0, None)], 4, S.HLLCPLUS)  # dim 4 with an adiabatic solver


SHAPES = [(112, 256), (96, 256), (64, 256), (48, 256), (32, 256), (48, 192), (32, 192), (32, 128)]


@pytest.mark.parametrize("kind,args", [("hex", (5,)), ("quad", (5,))])
def test_patch_shape_and_autotuner_leave_results_unchanged(kind, args):
    """Each edge's flux is evaluated from its half-edge of lower mesh index
    with the same inputs whatever patch holds it, and the same code handles
    every block size, so the state depends neither on the patch face cap nor
    on the threads per workgroup: fixed shapes (one and two face waves per
    patch), the default one and the tuned one agree bit for bit."""
    import splash_amd as S
    m = S.icosphere_hex(*args) if kind == "hex" else S.quad_sphere(*args)
    U0 = S.ic.adiabatic_state(m.face_centres(), GAM, seed=4, omega=(0, 0, 0.3))
    out = []
    shapes = [(0, 0), (64, 0), (112, 256), (128, 256), (32, 256), (17, 256), (48, 192), (100, 192), (40, 192),
              (32, 128), (20, 128), (-1, 0)]
    for cap, threads in shapes:
        p = S.default_params(patch_faces=cap, block_threads=threads)
        s = S.MUSCL_HLLCplus(m, U0, 5, GAM, 0.0, params=p, precision=S.MIXED)
        faces, nt = s.patch_shape()
        if cap > 0:
            assert faces == cap and nt == (threads or 256)
        elif cap == 0:
            assert (faces, nt) == ((112 if kind == "quad" else 64), 256)
        else:
            assert (faces, nt) in SHAPES
        s.do_steps(1e-4, 40)
        out.append(s.get_state())
        print(f"patch_faces {cap}, block_threads {threads} -> {faces}, {nt}")
    for U in out[1:]:
        assert np.array_equal(U, out[0])


def test_block_threads_rejects_other_sizes():
    import splash_amd as S
    m = S.icosphere_hex(2)
    U0 = S.ic.adiabatic_state(m.face_centres(), GAM, seed=4)
    with pytest.raises(S.SplashError):
        S.MUSCL_HLLCplus(m, U0, 5, GAM, 0.0, params=S.default_params(block_threads=64), precision=S.MIXED)


def test_ensemble_autotuned_equals_single_runs():
    import splash_amd as S
    runs = _runs(S)[:3]
    tuned = S.default_params(patch_faces=-1)
    ens = S.Ensemble([(m, U0, g, om, p if p is not None else tuned) for m, U0, g, om, p, _ in runs], 5,
                     S.HLLCPLUS)
    print("ensemble patch shape:", ens.patch_shape())
    assert ens.patch_shape() in SHAPES
    dt0 = [r[5] for r in runs]
    ens.do_steps(dt0, 20)
    for r, (m, U0, g, om, p, d) in enumerate(runs):
        s = S.MUSCL_HLLCplus(m, U0, 5, g, om, p is not None, params=p, precision=S.MIXED)
        s.do_steps(d, 20)
        assert np.array_equal(s.get_state(), ens.get_state(r))


def test_c4_full_ensemble():
    """C4 itself: 64 icosphere_hex runs of levels 4-7 (2.67M faces), HLLC+,
    500 steps in one ensemble.  Every run stays within 1e-5 relative L1
    density error of the same run advanced alone on the FP64 path, one run
    of each level equals a single FP32-flux solver bit for bit, and every
    run's time and stop_check match its FP64 run's."""
    import splash_amd as S
    import workloads as W
    runs = W.c4_runs(S)
    p = S.default_params(setup_threads=16)
    ens = S.Ensemble([(m, U0, GAM, 0.0, p) for m, U0, _ in runs], 5, S.HLLCPLUS)
    n_runs, faces, _, _ = ens.info()
    assert n_runs == 64 and faces == sum(m.n_faces for m, _, _ in runs)
    dt0 = [1e-4] * n_runs
    ens.do_steps(dt0, 500)
    worst = 0.0
    for r, (m, U0, lv) in enumerate(runs):
        f64 = S.MUSCL_HLLCplus(m, U0, 5, GAM, 0.0, params=p, precision=S.FP64)
        f64.do_steps(1e-4, 500)
        A, B = f64.get_state(), ens.get_state(r)
        l1 = float(np.sum(np.abs(B[:, 0] - A[:, 0])) / np.sum(np.abs(A[:, 0])))
        worst = max(worst, l1)
        assert l1 <= 1e-5, (r, lv, l1)
        assert ens.time(r) == pytest.approx(f64.time(), rel=1e-12)
        assert ens.get_stop_check(r) == f64.get_stop_check()
        del f64
    print(f"C4: worst per-run L1 density error vs FP64 after 500 steps {worst:.3e}")
    first_of_level = {}
    for r, (_, _, lv) in enumerate(runs):
        first_of_level.setdefault(lv, r)
    for lv, r in sorted(first_of_level.items()):
        m, U0, _ = runs[r]
        mx = S.MUSCL_HLLCplus(m, U0, 5, GAM, 0.0, params=p, precision=S.MIXED)
        mx.do_steps(1e-4, 500)
        assert np.array_equal(mx.get_state(), ens.get_state(r)), (r, lv)
        assert mx.time() == ens.time(r)
