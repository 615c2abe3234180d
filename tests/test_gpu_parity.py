"""FP64 HIP path vs the oracle (CPU restatement of the reference step).

Bar: the plain FP64 GPU state matches the oracle on identical inputs to
<= 1e-10 relative (per state component, normalised by that component's
largest magnitude) after the stated number of steps.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

GAM = 2 - 1 / (4 / 3)
TOL = 1e-10


def _rel(Ug, Ur):
    return float(np.max(np.max(np.abs(Ug - Ur), axis=0) / np.max(np.abs(Ur), axis=0)))


def _run_pair(S, kind, args, solver_cls, osolver, U0, dim, steps, dt, omega_ns=0.0, params=None,
              oparams=None, accretion_on=False):
    gen = {"hex": S.icosphere_hex, "quad": S.quad_sphere, "uv": S.uv_sphere, "ico": S.icosphere}[kind]
    mesh = gen(*args)
    if dim == 4:
        gpu = solver_cls(mesh, U0, 4, GAM, params=params)
    else:
        gpu = solver_cls(mesh, U0, 5, GAM, omega_ns, accretion_on, params=params)
    om = O.Mesh(kind, *args)
    ref = O.Solver(om, U0, dim, osolver, GAM, omega_ns, oparams, threads=8)
    for _ in range(steps):
        gpu.do_step(dt)
        ref.do_step(dt)
    return gpu, ref


@pytest.mark.parametrize("solver", ["HLLCplus", "HLLC", "HLLE_p"])
def test_adiabatic_fp64_matches_oracle(solver):
    import splash_amd as S
    cls = getattr(S, "MUSCL_" + solver)
    osolver = {"HLLCplus": O.HLLCPLUS, "HLLC": O.HLLC, "HLLE_p": O.HLLE_P}[solver]
    mesh = S.icosphere_hex(4)
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=2, omega=(0, 0, 0.3))
    gpu, ref = _run_pair(S, "hex", (4,), cls, osolver, U0, 5, 50, 1e-4)
    err = _rel(gpu.get_state(), ref.state())
    print(f"{solver} hex(4) 50 steps: max rel err {err:.3e}")
    assert err <= TOL
    assert abs(gpu.time() - ref.time()) <= 1e-15 * ref.time()
    assert gpu.get_stop_check() == ref.get_stop_check()
