"""HIP path vs the oracle (CPU restatement of the reference step).

Bars:
- plain FP64 GPU state matches the oracle on identical inputs to <= 1e-10
  relative (per state component, normalised by that component's largest
  magnitude) after the stated number of steps; time and stop_check agree;
- the FP32-flux / FP64-state path stays within 1e-5 relative L1 density
  error of the FP64 GPU path after 1000 steps and conserves mass to 1e-13
  relative with sources off;
- graph-replayed steps equal eager steps bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as O
import workloads as W
from lightcurve_ref import light_curves

pytestmark = pytest.mark.gpu

GAM = 2 - 1 / (4 / 3)
TOL = 1e-10


def _rel(Ug, Ur):
    return float(np.max(np.max(np.abs(Ug - Ur), axis=0) / np.max(np.abs(Ur), axis=0)))


def _gen(S, kind):
    return {"hex": S.icosphere_hex, "quad": S.quad_sphere, "uv": S.uv_sphere, "ico": S.icosphere}[kind]


def _oparams(p):
    return O.make_params(accretion_on=bool(p.accretion_on), friction_on=bool(p.friction_on), acc_rate=p.acc_rate,
                         e_acc=p.e_acc, omega_acc_abs=p.omega_acc_abs, tilt_angle=p.tilt_angle,
                         acc_width=p.acc_width, alpha=p.alpha)


def _pair(S, kind, args, solver, U0, steps, dt, omega_ns=0.0, params=None, threads=8):
    mesh = _gen(S, kind)(*args)
    params = params if params is not None else S.default_params()
    osolver = {"HLLCplus": O.HLLCPLUS, "HLLC": O.HLLC, "HLLE_p": O.HLLE_P, "HLLE": O.HLLE}[solver]
    cls = getattr(S, "MUSCL_" + solver)
    if solver == "HLLE":
        gpu = cls(mesh, U0, 4, GAM, params=params)
        dim = 4
    else:
        gpu = cls(mesh, U0, 5, GAM, omega_ns, bool(params.accretion_on), params=params)
        dim = 5
    ref = O.Solver(O.Mesh(kind, *args), U0, dim, osolver, GAM, omega_ns, _oparams(params), threads=threads)
    for _ in range(steps):
        gpu.do_step(dt)
        ref.do_step(dt)
    return gpu, ref


@pytest.mark.parametrize("solver", ["HLLCplus", "HLLC", "HLLE_p"])
def test_adiabatic_fp64_matches_oracle(solver):
    import splash_amd as S
    mesh = S.icosphere_hex(4)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=2, omega=(0, 0, 0.3))
    gpu, ref = _pair(S, "hex", (4,), solver, U0, 50, 1e-4)
    err = _rel(gpu.get_state(), ref.state())
    print(f"{solver} hex(4) 50 steps: max rel err {err:.3e}")
    assert err <= TOL
    assert abs(gpu.time() - ref.time()) <= 1e-15 * ref.time()
    assert gpu.last_dt() == pytest.approx(ref.last_dt(), rel=1e-14)
    assert gpu.get_stop_check() == ref.get_stop_check()


@pytest.mark.parametrize("kind,args", [("quad", (4,)), ("uv", (32, 16)), ("ico", (3,))])
def test_other_meshes_fp64_match_oracle(kind, args):
    import splash_amd as S
    mesh = _gen(S, kind)(*args)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=7, omega=(0.1, 0.2, 0.3))
    gpu, ref = _pair(S, kind, args, "HLLCplus", U0, 30, 1e-4)
    err = _rel(gpu.get_state(), ref.state())
    print(f"HLLC+ {kind}{args} 30 steps: max rel err {err:.3e}")
    assert err <= TOL


def test_c1_isothermal_hlle_matches_oracle():
    """C1 at its own size: MUSCL_HLLE, dim 4, icosphere_hex(5) (10,242 faces),
    Sigma = 1 + 0.05 N at rest, 100 steps of dt 1e-3."""
    import splash_amd as S
    kind, args, solver, _, _, state, steps, dt = W.WORKLOADS["c1"]
    mesh = W.make_mesh(S, kind, args)
    U0 = W.initial_state(S, mesh, state)
    gpu, ref = _pair(S, kind, args, solver, U0, steps, dt, threads=16)
    err = _rel(gpu.get_state(), ref.state())
    print(f"C1 HLLE dim4 hex(5) {steps} steps: max rel err {err:.3e}")
    assert err <= TOL
    assert gpu.time() == pytest.approx(ref.time(), rel=1e-14)
    assert gpu.get_stop_check() == ref.get_stop_check()


@pytest.mark.parametrize("friction", [True, False])
def test_accretion_sources_match_oracle(friction):
    import splash_amd as S
    p = S.default_params(friction_on=int(friction), accretion_on=1, acc_rate=1e-3, e_acc=0.1,
                         omega_acc_abs=0.8, tilt_angle=0.3, acc_width=0.1, alpha=1e-6)
    mesh = S.quad_sphere(4)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=3, omega=(0, 0, 0.1))
    gpu, ref = _pair(S, "quad", (4,), "HLLC", U0, 40, 1e-4, omega_ns=0.1, params=p)
    err = _rel(gpu.get_state(), ref.state())
    print(f"HLLC accretion friction={friction} quad(4) 40 steps: max rel err {err:.3e}")
    assert err <= TOL
    assert np.max(np.abs(gpu.get_state()[:, 0] - U0[:, 0])) > 1e-8  # sources acted


def test_stop_check_and_negative_dt_match_oracle():
    """Omega = 0.5 makes g_eff < 0: the friction dt limit goes negative and
    stop_check is raised on step 1, as on the reference build."""
    import splash_amd as S
    mesh = S.icosphere_hex(4)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=2, omega=(0, 0, 0.5))
    gpu, ref = _pair(S, "hex", (4,), "HLLCplus", U0, 1, 1e-4)
    assert ref.get_stop_check() and gpu.get_stop_check()
    bits, _ = gpu.stop_bits()
    assert bits & 8


def test_mixed_dt_limits_match_fp64_and_oracle():
    """The mixed path forms dt inside its stage-1 kernel from limits its
    stage 2 (or the prepare kernel) folded into accumulator keys.  With the
    dt clamp active (dt0 = 1 on hex(5)) its first dt equals the FP64
    path's and the oracle's, both from the same initial state (the mixed
    path's |v| and c_s differ only by rounding), at every later step it
    stays clamped, and a negative friction limit (Omega = 0.5) raises the
    same stop bit without clamping dt, as on the reference build."""
    import splash_amd as S
    mesh = S.icosphere_hex(5)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=3, omega=(0, 0, 0.3))
    mixed = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=S.MIXED)
    gpu, ref = _pair(S, "hex", (5,), "HLLCplus", U0, 1, 1.0)
    mixed.do_step(1.0)
    assert ref.time() < 1e-2 and gpu.time() == pytest.approx(ref.time(), rel=1e-14)
    assert mixed.time() == pytest.approx(gpu.time(), rel=1e-13)
    assert not mixed.get_stop_check()
    for _ in range(4):
        mixed.do_step(1.0)
        assert mixed.last_dt() < 1e-2
    U1 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=2, omega=(0, 0, 0.5))
    mixed = S.MUSCL_HLLCplus(mesh, U1, 5, GAM, 0.0, precision=S.MIXED)
    gpu, ref = _pair(S, "hex", (5,), "HLLCplus", U1, 1, 1e-4)
    mixed.do_step(1e-4)
    assert ref.get_stop_check() and gpu.get_stop_check() and mixed.get_stop_check()
    assert mixed.stop_bits()[0] & 8
    assert mixed.time() == pytest.approx(ref.time(), rel=1e-13)


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_graph_replay_equals_eager(precision):
    """Graph replays equal eager steps; for the mixed path the replays also
    carry its dt (formed in stage 1 from the accumulator buffer the previous
    stage 2 filled, the buffer chosen from the device step count), so any
    mix of replays and eager steps must keep the buffers in step.  dt0
    changes between calls (the replays read it from a device scalar): 1.0
    leaves dt to the CFL clamp, 5e-5 sets it."""
    import splash_amd as S
    mesh = S.icosphere_hex(4)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=5, omega=(0, 0, 0.2))
    prec = getattr(S, precision)
    a = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=prec)
    b = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=prec)
    for dt0 in (1.0, 5e-5, 1.0, 5e-5):
        a.do_steps(dt0, 19)  # two 8-step graph replays + 3 eager steps each time
        for _ in range(19):
            b.do_step(dt0)
        assert np.array_equal(a.get_state(), b.get_state())
        assert a.time() == b.time() and a.last_dt() == b.last_dt()
        assert (a.last_dt() < 1e-2) if dt0 == 1.0 else (a.last_dt() == dt0)


def test_mixed_precision_error_and_conservation():
    import splash_amd as S
    mesh = S.icosphere_hex(5)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=2, omega=(0, 0, 0.3))
    f64 = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=S.FP64)
    mx = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=S.MIXED)
    d0 = mx.diagnostics()
    f64.do_steps(1e-4, 1000)
    mx.do_steps(1e-4, 1000)
    A, B = f64.get_state(), mx.get_state()
    l1 = np.sum(np.abs(A[:, 0] - B[:, 0])) / np.sum(np.abs(A[:, 0]))
    print(f"mixed vs fp64, hex(5) 1000 steps: L1 rel density err {l1:.3e}")
    assert l1 <= 1e-5
    d1 = mx.diagnostics()
    # each edge's flux is evaluated once per patch from identical inputs and
    # applied as +F and -F to its two faces, so mass changes only by the FP64
    # rounding of each face's update: round-off after 1000 steps
    assert abs(d1[0] / d0[0] - 1) < 1e-13
    assert mx.time() == f64.time()


@pytest.mark.parametrize("solver", ["HLLC", "HLLE"])
def test_mixed_density_floor_near_vacuum(solver):
    """Twelve faces start below the density floor (Sigma = 1e-9, at rest;
    adiabatic: p = 1e-3) in a rotating disc: both paths floor Sigma at 1e-8
    in every stage (the FP32-flux path's stage 1 on its FP32 stage input,
    stage 2 on the FP64 state), stay finite, raise the same stop_check (the
    adiabatic near-vacuum faces' speeds drive the effective gravity
    negative, as on the reference), and agree to the L1 bar over the mesh."""
    import splash_amd as S
    mesh = S.icosphere_hex(4)
    fc = mesh.face_centres()
    c = fc[0] / np.linalg.norm(fc[0])
    near = np.argsort(-(fc @ c))[:12]
    if solver == "HLLC":
        U0 = S.ic.adiabatic_state(fc, GAM, seed=3, omega=(0, 0, 0.1))
        U0[near, 4] = 1e-3 / (GAM - 1)
        make = lambda prec: S.MUSCL_HLLC(mesh, U0, 5, GAM, 0.0, precision=prec)  # noqa: E731
    else:
        U0 = S.ic.isothermal_state(fc, seed=3, sigma_amp=0.05)
        make = lambda prec: S.MUSCL_HLLE(mesh, U0, 4, GAM, precision=prec)  # noqa: E731
    U0[near, 0] = 1e-9
    U0[near, 1:4] = 0
    f64, mx = make(S.FP64), make(S.MIXED)
    for s in (f64, mx):
        s.do_steps(1e-4, 20)
    A, B = f64.get_state(), mx.get_state()
    assert np.isfinite(A).all() and np.isfinite(B).all()
    assert A[:, 0].min() >= 1e-8 and B[:, 0].min() >= 1e-8
    assert mx.get_stop_check() == f64.get_stop_check()
    l1 = np.sum(np.abs(A[:, 0] - B[:, 0])) / np.sum(np.abs(A[:, 0]))
    print(f"near-vacuum {solver} hex(4) 20 steps: min Sigma {A[:, 0].min():.3e} / {B[:, 0].min():.3e}, "
          f"stop {f64.get_stop_check()}, L1 {l1:.3e}")
    assert l1 <= 1e-5


@pytest.mark.parametrize("kind,args,solver", [("hex", (5,), "HLLC"), ("quad", (6,), "HLLCplus")])
def test_mixed_discontinuous_state(kind, args, solver):
    """A Sod-like jump at rest (Sigma 1 / 0.125, p 1 / 0.1 across the
    z = 0.2 circle) for 200 steps: the shock, contact and rarefaction run
    through the limiter's switches, and the FP32-flux path stays finite,
    keeps the FP64 path's time and stop_check, and meets the L1 bar."""
    import splash_amd as S
    mesh = S.icosphere_hex(*args) if kind == "hex" else S.quad_sphere(*args)
    fc = mesh.face_centres()
    z = fc[:, 2] / np.linalg.norm(fc, axis=1)
    U0 = np.zeros((mesh.n_faces, 5))
    U0[:, 0] = np.where(z > 0.2, 1.0, 0.125)
    U0[:, 4] = np.where(z > 0.2, 1.0, 0.1) / (GAM - 1)
    cls = getattr(S, "MUSCL_" + solver)
    f64 = cls(mesh, U0, 5, GAM, 0.0, precision=S.FP64)
    mx = cls(mesh, U0, 5, GAM, 0.0, precision=S.MIXED)
    for s in (f64, mx):
        s.do_steps(1e-4, 200)
    A, B = f64.get_state(), mx.get_state()
    assert np.isfinite(B).all()
    assert mx.time() == f64.time() and mx.get_stop_check() == f64.get_stop_check()
    l1 = np.sum(np.abs(A[:, 0] - B[:, 0])) / np.sum(np.abs(A[:, 0]))
    print(f"discontinuous {kind}{args} {solver} 200 steps: L1 rel density err {l1:.3e}")
    assert l1 <= 1e-5


def test_mixed_fast_rotation():
    """Solid rotation at Omega = 1.5 (|v| up to 1.5, kinetic energy up to 30
    times the pressure): most of stage 1's faces are formed from the FP64
    state rather than by the FP32 pressure update, and the FP32-flux path
    keeps the FP64 path's time and stop_check and meets the L1 bar."""
    import splash_amd as S
    mesh = S.quad_sphere(5)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=6, omega=(0, 0, 1.5))
    f64 = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=S.FP64)
    mx = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=S.MIXED)
    for s in (f64, mx):
        s.do_steps(1e-4, 100)
    A, B = f64.get_state(), mx.get_state()
    assert np.isfinite(B).all()
    assert mx.time() == f64.time() and mx.get_stop_check() == f64.get_stop_check()
    l1 = np.sum(np.abs(A[:, 0] - B[:, 0])) / np.sum(np.abs(A[:, 0]))
    print(f"fast rotation quad(5) 100 steps: L1 rel density err {l1:.3e}, stop {f64.get_stop_check()}")
    assert l1 <= 1e-5


def test_c2_full_size_fp64_parity_and_conservation():
    """icosphere_hex(7) (163,842 faces): FP64 GPU vs oracle after 100 steps
    of the C2 workload, plus mass/energy/angular-momentum conservation."""
    import splash_amd as S
    mesh = S.icosphere_hex(7)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=2, omega=(0, 0, 0.5))
    gpu, ref = _pair(S, "hex", (7,), "HLLCplus", U0, 100, 1e-4, threads=16)
    err = _rel(gpu.get_state(), ref.state())
    print(f"C2 hex(7) 100 steps: max rel err {err:.3e}")
    assert err <= TOL
    d0 = ref.diagnostics()
    d = gpu.diagnostics()
    assert np.allclose(d, d0, rtol=1e-13, atol=1e-15)


def test_c3_full_size_fp64_parity_and_mixed_error():
    """C3 at full size: quad_sphere(9) (1,572,864 faces), HLLC, omega_ns =
    0.1, accretion on (rate 1e-3, e_acc 0.1, omega_acc 0.8, tilt 0.3, width
    0.1) and friction with alpha = 1e-6, the benchmark's state.  The FP64
    path matches the oracle to 1e-10 after 100 steps, sources included, and
    the FP32-flux path stays within 1e-5 relative L1 density error of the
    FP64 path after 1000 steps."""
    import splash_amd as S
    kind, args, solver, omega_ns, _, state, _, dt = W.WORKLOADS["c3"]
    p = S.default_params(setup_threads=16, **W.ACCRETION)
    mesh = W.make_mesh(S, kind, args)
    U0 = W.initial_state(S, mesh, state)
    gpu, ref = _pair(S, kind, args, solver, U0, 100, dt, omega_ns=omega_ns, params=p, threads=16)
    err = _rel(gpu.get_state(), ref.state())
    print(f"C3 quad(9) 100 steps: FP64 max rel err vs oracle {err:.3e}, dt {gpu.last_dt():.4e}")
    assert err <= TOL
    assert gpu.time() == pytest.approx(ref.time(), rel=1e-12)
    assert not gpu.get_stop_check() and not ref.get_stop_check()
    del gpu, ref
    f64 = S.MUSCL_HLLC(mesh, U0, 5, GAM, omega_ns, True, params=p, precision=S.FP64)
    mx = S.MUSCL_HLLC(mesh, U0, 5, GAM, omega_ns, True, params=p, precision=S.MIXED)
    f64.do_steps(dt, 1000)
    mx.do_steps(dt, 1000)
    A, B = f64.get_state(), mx.get_state()
    l1 = np.sum(np.abs(B[:, 0] - A[:, 0])) / np.sum(np.abs(A[:, 0]))
    print(f"C3 quad(9) 1000 steps: FP32-flux vs FP64 L1 rel density err {l1:.3e}")
    assert l1 <= 1e-5


@pytest.mark.parametrize("name", ["c5a", "c5b"])
def test_c5_full_size_fp64_parity_and_mixed_error(name):
    """C5a uv_sphere(2048, 1024) (2,097,152 faces: quads plus the pole
    triangle fans) and C5b icosphere_hex(9) (2,621,442 faces) with their own
    physics: HLLC+, omega_ns = 0.1, accretion and friction as C3, a
    spherical-harmonic density at rest.  The FP64 path matches the oracle to
    1e-10 after 100 steps (C5a) / 20 steps (C5b, whose oracle run is the
    slowest), and the FP32-flux path stays within 1e-5 relative L1 density
    error of the FP64 path after 1000 steps.  (The flow starts at rest, so
    the momenta stay small: the bar is per component, against that
    component's largest magnitude.)"""
    import splash_amd as S
    kind, args, solver, omega_ns, _, state, _, dt = W.WORKLOADS[name]
    p = S.default_params(setup_threads=16, **W.ACCRETION)
    mesh = W.make_mesh(S, kind, args)
    U0 = W.initial_state(S, mesh, state)
    n = 100 if name == "c5a" else 20
    gpu, ref = _pair(S, kind, args, solver, U0, n, dt, omega_ns=omega_ns, params=p, threads=16)
    err = _rel(gpu.get_state(), ref.state())
    print(f"{name} {kind}{args} {n} steps: FP64 max rel err vs oracle {err:.3e}, dt {gpu.last_dt():.3e}")
    assert err <= TOL
    assert gpu.time() == pytest.approx(ref.time(), rel=1e-12)
    del gpu, ref
    f64 = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, omega_ns, True, params=p, precision=S.FP64)
    mx = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, omega_ns, True, params=p, precision=S.MIXED)
    f64.do_steps(dt, 1000)
    mx.do_steps(dt, 1000)
    A, B = f64.get_state(), mx.get_state()
    l1 = np.sum(np.abs(B[:, 0] - A[:, 0])) / np.sum(np.abs(A[:, 0]))
    print(f"{name} 1000 steps: FP32-flux vs FP64 L1 rel density err {l1:.3e}")
    assert l1 <= 1e-5
    assert mx.time() == pytest.approx(f64.time(), rel=1e-12)


def test_hex9_mixed_vs_fp64_and_conservation():
    """icosphere_hex(9) (2,621,442 faces, 15.7M half-edges), the largest
    single-mesh workload: after 100 steps of HLLC+ with sources off, the
    FP32-flux path stays within 1e-5 relative L1 density of the FP64 path,
    both conserve mass to round-off, and both report the same time and
    stop_check.  (The oracle would take minutes here; size-independent
    properties stand in for it.)"""
    import splash_amd as S
    mesh = S.icosphere_hex(9)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=5, omega=(0.2, 0.0, 0.4))
    p = S.default_params(setup_threads=16)
    a = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, params=p, precision=S.FP64)
    b = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, params=p, precision=S.MIXED)
    d0 = a.diagnostics()
    a.do_steps(1e-4, 100)
    b.do_steps(1e-4, 100)
    Ua, Ub = a.get_state(), b.get_state()
    l1 = np.sum(np.abs(Ub[:, 0] - Ua[:, 0])) / np.sum(np.abs(Ua[:, 0]))
    print(f"C5b hex(9) 100 steps: mixed vs FP64 L1 {l1:.3e}, time {a.time():.6e}")
    assert l1 <= 1e-5
    assert a.time() == pytest.approx(b.time(), rel=1e-12) and a.get_stop_check() == b.get_stop_check()
    for s in (a, b):
        d = s.diagnostics()
        assert abs(d[0] / d0[0] - 1) < 1e-12


def test_hex9_fp64_parity_with_split_dt_reduction():
    """icosphere_hex(9): about 61k FP64 patches, so the dt kernel splits each
    step's partials over 15 workgroups and the last one finishes dt.  With
    dt0 = 1e-3 the CFL clamp is active, so dt is exactly that reduction's
    result: after 2 steps the FP64 state matches the oracle to 1e-10 and dt
    and time to round-off; the FP32-flux path's first dt (its stage 1's
    reduction of the limits its prepare kernel folded into the accumulator
    slots) agrees too."""
    import splash_amd as S
    mesh = S.icosphere_hex(9)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=8, omega=(0.1, 0.0, 0.3))
    p = S.default_params(setup_threads=16)
    gpu = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, params=p, precision=S.FP64)
    mixed = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, params=p, precision=S.MIXED)
    ref = O.Solver(O.Mesh("hex", 9), U0, 5, O.HLLCPLUS, GAM, 0.0, _oparams(p), threads=16)
    ref.do_step(1e-3)
    gpu.do_step(1e-3)
    mixed.do_step(1e-3)
    dt1 = ref.last_dt()
    assert dt1 < 1e-3  # the CFL / friction clamp decided dt
    assert gpu.last_dt() == pytest.approx(dt1, rel=1e-14)
    assert mixed.last_dt() == pytest.approx(dt1, rel=1e-12)
    ref.do_step(1e-3)
    gpu.do_step(1e-3)
    err = _rel(gpu.get_state(), ref.state())
    print(f"hex(9) FP64 2 steps: max rel err {err:.3e}, dt {dt1:.6e} / {ref.last_dt():.6e}")
    assert err <= TOL
    assert gpu.last_dt() == pytest.approx(ref.last_dt(), rel=1e-14)
    assert gpu.time() == pytest.approx(ref.time(), rel=1e-14)


def test_profile_step_times_back_to_back_steps():
    """profile_step advances exactly n steps, gives positive per-kernel means
    that add up to the step time, and leaves the state where as many plain
    steps leave it.  The mixed path launches no dt kernel (its stage 1 forms
    dt), so its first mean is 0."""
    import splash_amd as S
    mesh = S.icosphere_hex(4)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=5, omega=(0, 0, 0.2))
    for prec in (S.FP64, S.MIXED):
        a = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=prec)
        b = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=prec)
        ms = a.profile_step(1e-4, 5)
        for _ in range(5):
            b.do_step(1e-4)
        assert a.time() == b.time()
        assert np.array_equal(a.get_state(), b.get_state())
        launched = ms[1:] if prec == S.MIXED else ms
        assert np.all(launched > 0) and ms[:3].sum() <= ms[3] * 1.0001
        if prec == S.MIXED:
            assert ms[0] == 0


@pytest.mark.parametrize("solver,kind,args,sources", [
    ("HLLE", "hex", (4,), False),
    ("HLLE_p", "hex", (4,), False),
    ("HLLC", "quad", (5,), True),
    ("HLLCplus", "uv", (64, 32), False),
    ("HLLCplus", "ico", (4,), True),
])
def test_mixed_precision_every_solver_and_mesh(solver, kind, args, sources):
    """The FP32-flux path against the FP64 path after 1000 steps on every
    solver, on triangle / quad / hexagon meshes, with and without sources:
    <= 1e-5 relative L1 density error, the same time, and (sources off) mass
    conserved to round-off."""
    import splash_amd as S
    mesh = _gen(S, kind)(*args)
    cls = getattr(S, "MUSCL_" + solver)
    p = (S.default_params(friction_on=1, accretion_on=1, acc_rate=1e-3, e_acc=0.1, omega_acc_abs=0.8,
                          tilt_angle=0.3, acc_width=0.1, alpha=1e-6) if sources else S.default_params())
    if solver == "HLLE":
        U0 = S.ic.isothermal_state(mesh.face_centres(), seed=1, sigma_amp=0.05)
        make = lambda prec: cls(mesh, U0, 4, GAM, params=p, precision=prec)  # noqa: E731
        dt = 1e-3
    else:
        U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=3, omega=(0.1, 0.0, 0.3))
        make = lambda prec: cls(mesh, U0, 5, GAM, 0.1, sources, params=p, precision=prec)  # noqa: E731
        dt = 1e-4
    f64, mx = make(S.FP64), make(S.MIXED)
    d0 = mx.diagnostics()
    f64.do_steps(dt, 1000)
    mx.do_steps(dt, 1000)
    A, B = f64.get_state(), mx.get_state()
    assert np.all(np.isfinite(B))
    l1 = np.sum(np.abs(A[:, 0] - B[:, 0])) / np.sum(np.abs(A[:, 0]))
    print(f"mixed vs fp64, {solver} {kind}{args} sources={sources}: L1 rel density err {l1:.3e}")
    assert l1 <= 1e-5
    assert mx.time() == pytest.approx(f64.time(), rel=1e-12)
    if not sources:
        assert abs(mx.diagnostics()[0] / d0[0] - 1) < 1e-13


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_light_curves_device_reduction(precision):
    import splash_amd as S
    mesh = S.icosphere_hex(5)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=6, omega=(0.1, 0, 0.3))
    s = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=getattr(S, precision))
    s.do_steps(1e-4, 30)
    g = S.Geometry(mesh)
    want = light_curves(s.get_state(), g.fc, g.area, GAM)
    got = s.get_light_curves()
    print(f"light curves {precision}: {got}")
    assert np.all(got > 0)
    np.testing.assert_allclose(got, want, rtol=1e-12)


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_driver_readout_is_one_per_state(precision):
    """The diagnostics sums, light curves and scalars come from one device
    readout per state: repeated queries agree bit for bit, a step replaces
    them (the light curves then match the stepped state), and after set_state
    they equal those of a fresh solver built on the same state."""
    import splash_amd as S
    mesh = S.icosphere_hex(4)
    prec = getattr(S, precision)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=8, omega=(0, 0, 0.2))
    U1 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=9, omega=(0.1, 0, 0.3))
    a = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=prec)
    d0, l0 = a.diagnostics(), a.get_light_curves()
    assert np.array_equal(a.diagnostics(), d0) and np.array_equal(a.get_light_curves(), l0)
    a.do_step(1e-4)
    assert a.time() > 0
    g = S.Geometry(mesh)
    l1 = a.get_light_curves()
    assert not np.array_equal(l1, l0)
    np.testing.assert_allclose(l1, light_curves(a.get_state(), g.fc, g.area, GAM), rtol=1e-12)
    a.set_state(U1, 0.5, 0)
    b = S.MUSCL_HLLCplus(mesh, U1, 5, GAM, 0.0, precision=prec)
    assert np.array_equal(a.diagnostics(), b.diagnostics())
    assert np.array_equal(a.get_light_curves(), b.get_light_curves())
    assert a.time() == 0.5 and b.time() == 0.0


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_set_state_resumes_exactly(precision):
    """Resuming from (state, time, stop bits) continues bit for bit as the
    run that produced them."""
    import splash_amd as S
    mesh = S.quad_sphere(4)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=9, omega=(0, 0.2, 0.2))
    prec = getattr(S, precision)
    a = S.MUSCL_HLLC(mesh, U0, 5, GAM, 0.0, precision=prec)
    a.do_steps(1e-4, 13)
    U, t = a.get_state(), a.time()
    bits, _ = a.stop_bits()
    a.do_steps(1e-4, 11)
    b = S.MUSCL_HLLC(mesh, U0, 5, GAM, 0.0, precision=prec)
    b.set_state(U, t, bits)
    b.do_steps(1e-4, 11)
    assert np.array_equal(a.get_state(), b.get_state())
    assert a.time() == b.time() and a.get_stop_check() == b.get_stop_check()


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
@pytest.mark.parametrize("column", [0, 4])
def test_nan_state_stop_check_matches_oracle(precision, column):
    """A NaN density reaches its neighbours' fluxes on the first step and
    raises stop_check (MUSCL_base.hpp:264-268); a NaN energy does not: the
    pressure floor max(1e-16, p) (adiabatic.hpp:582-592) turns it into 1e-16.
    Both paths agree with the oracle on which."""
    import splash_amd as S
    mesh = S.icosphere_hex(3)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=2, omega=(0, 0, 0.2))
    U0[17, column] = np.nan
    s = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, precision=getattr(S, precision))
    s.do_step(1e-4)
    ref = O.Solver(O.Mesh("hex", 3), U0, 5, O.HLLCPLUS, GAM, 0.0, _oparams(S.default_params()))
    ref.do_step(1e-4)
    assert s.get_stop_check() == ref.get_stop_check() == (column == 0)
    if column == 0:
        bits, _ = s.stop_bits()
        assert bits & (1 | 2)


@pytest.mark.parametrize("solver,kind,args", [("HLLCplus", "hex", (4,)), ("HLLC", "quad", (4,)),
                                              ("HLLE", "hex", (4,))])
def test_curl_matches_oracle(solver, kind, args):
    """write_t_curl's per-face values: zero before a step and after set_state
    (the reference's freshly sized U_plus); after steps the FP64 path matches
    the oracle to 1e-10 of the largest magnitude and the FP32-flux path to
    2e-4 of it."""
    import splash_amd as S
    mesh = _gen(S, kind)(*args)
    fc = S.ic.unit_centres(mesh.face_centres())
    if solver == "HLLE":
        U0 = S.ic.isothermal_state(mesh.face_centres(), seed=3)
        v = np.cross([0.1, 0.2, 0.3], fc)
        U0[:, 1:4] = U0[:, :1] * np.cross(fc, v)
    else:
        U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=3, omega=(0.1, 0.2, 0.3))
    gpu, ref = _pair(S, kind, args, solver, U0, 0, 1e-4)
    assert np.array_equal(gpu.curl(), np.zeros(mesh.n_faces))
    for _ in range(20):
        gpu.do_step(1e-4)
        ref.do_step(1e-4)
    want = ref.curl()
    scale = np.max(np.abs(want))
    err = np.max(np.abs(gpu.curl() - want)) / scale
    cls = getattr(S, "MUSCL_" + solver)
    if solver == "HLLE":
        mixed = cls(mesh, U0, 4, GAM, precision=S.MIXED)
    else:
        mixed = cls(mesh, U0, 5, GAM, 0.0, False, precision=S.MIXED)
    mixed.do_steps(1e-4, 20)
    err32 = np.max(np.abs(mixed.curl() - want)) / scale
    print(f"curl {solver} {kind}{args}: |curl| max {scale:.3e}, FP64 err {err:.3e}, mixed err {err32:.3e}")
    assert err <= TOL
    assert err32 <= 2e-4
    gpu.set_state(gpu.get_state(), gpu.time())
    assert np.array_equal(gpu.curl(), np.zeros(mesh.n_faces))
