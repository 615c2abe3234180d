"""The benchmark configurations (C1-C5) as seeded synthetic workloads, shared
by bench.py and the GPU parity tests.

C1  icosphere_hex(5), dim 4 HLLE, Sigma = 1 + 0.05 N, at rest, dt 1e-3, 100 steps
C2  icosphere_hex(7), dim 5 HLLC+, Sigma = 1 + 0.1 N, Pi = 0.1 Sigma,
    Omega = 0.5 z + 0.01 N, dt 1e-4, 1000 steps
C3  quad_sphere(9), dim 5 HLLC, omega_ns 0.1, accretion on (rate 1e-3,
    e_acc 0.1, omega_acc 0.8, tilt 0.3, width 0.1, friction alpha 1e-6),
    Omega = 0.3 z + 0.01 N (Omega = 0.5 as in C2 trips stop_check on step 1),
    dt 1e-4, 1000 steps
C4  64 icosphere_hex runs, levels drawn from {4..7} with seed 4, run i seeded
    4000 + i, random rotation axis, Omega ~ U(0,1), dim 5 HLLC+, 500 steps
C5a uv_sphere(2048, 1024) and C5b icosphere_hex(9), dim 5 HLLC+, omega_ns 0.1,
    accretion as C3, Sigma = 1 + 8 random real Y_lm (l <= 16, amplitude 0.05),
    Pi = 0.1, at rest, 2000 steps
"""
import numpy as np

GAM = 2 - 1 / (4 / 3)
ACCRETION = dict(accretion_on=1, friction_on=1, acc_rate=1e-3, e_acc=0.1, omega_acc_abs=0.8, tilt_angle=0.3,
                 acc_width=0.1, alpha=1e-6)

# name: (mesh kind, mesh args, solver, omega_ns, accretion, initial state, steps, dt0)
WORKLOADS = {
    "c1": ("hex", (5,), "HLLE", 0.0, False, "isothermal", 100, 1e-3),
    "c2": ("hex", (7,), "HLLCplus", 0.0, False, "rotation05", 1000, 1e-4),
    "c3": ("quad", (9,), "HLLC", 0.1, True, "rotation03", 1000, 1e-4),
    "c5a": ("uv", (2048, 1024), "HLLCplus", 0.1, True, "harmonics", 2000, 1e-4),
    "c5b": ("hex", (9,), "HLLCplus", 0.1, True, "harmonics", 2000, 1e-4),
}
STATES = {
    "rotation03": "Sigma=1+0.1N, Pi=0.1Sigma, Omega=0.3 z + 0.01N",
    "rotation05": "Sigma=1+0.1N, Pi=0.1Sigma, Omega=0.5 z + 0.01N",
    "harmonics": "Sigma=1+8 real Y_lm (l<=16, amplitude 0.05), Pi=0.1, at rest",
    "isothermal": "Sigma=1+0.05N, at rest",
}


def make_mesh(S, kind, args):
    return {"hex": S.icosphere_hex, "quad": S.quad_sphere, "uv": S.uv_sphere, "ico": S.icosphere}[kind](*args)


def initial_state(S, mesh, state):
    fc = mesh.face_centres()
    if state == "isothermal":
        return S.ic.isothermal_state(fc, seed=1, sigma_amp=0.05)
    if state.startswith("rotation"):
        om = {"rotation03": 0.3, "rotation05": 0.5}[state]
        return S.ic.adiabatic_state(fc, GAM, seed=2, sigma_amp=0.1, pi_frac=0.1, omega=(0.0, 0.0, om), noise=0.01)
    sigma = 1.0 + S.ic.real_sph_harm_sum(fc, seed=5, n_modes=8, lmax=16, amplitude=0.05)
    return S.ic.adiabatic_state(fc, GAM, seed=5, sigma=sigma, pressure=np.full_like(sigma, 0.1), noise=0.0)


def c4_runs(S, n_runs=64, seed=4):
    """C4's runs as (mesh, U0, level)."""
    levels = 4 + np.minimum((S.ic.uniform(seed, 0, n_runs) * 4).astype(int), 3)
    meshes = {lv: S.icosphere_hex(int(lv)) for lv in sorted(set(levels.tolist()))}
    runs = []
    for i, lv in enumerate(levels.tolist()):
        si = 4000 + i
        axis = S.ic.gaussian(si, 7, 3)
        axis /= np.linalg.norm(axis)
        omega = S.ic.uniform(si, 8, 1)[0] * axis
        m = meshes[lv]
        U0 = S.ic.adiabatic_state(m.face_centres(), GAM, seed=si, sigma_amp=0.1, pi_frac=0.1,
                                  omega=tuple(omega), noise=0.01)
        runs.append((m, U0, lv))
    return runs
