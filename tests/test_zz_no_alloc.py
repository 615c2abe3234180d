"""splash_readout_noalloc, the allocation-free variant of the readout behind
splash_diagnostics / splash_light_curves: called through the C ABI into
device memory, it gives the same nine sums bit for bit, on consecutive calls
whose second and third runs are smaller than the first (the library's fixed
partial-sum array keeps the previous call's contents)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAM = 2 - 1 / (4 / 3)


class _DeviceBuffer:
    """Device memory for the output through the HIP runtime the library
    itself loaded (the test allocates; the entry point does not)."""

    def __init__(self, n):
        with open("/proc/self/maps") as maps:
            path = next((ln.split()[-1] for ln in maps if "libamdhip64" in ln), "libamdhip64.so.7")
        self.hip = C.CDLL(path)
        self.n = n
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0
        host = np.full(n, np.nan)
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(8 * n), 1) == 0

    def read(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(S, s, buf):
    from splash_amd._lib import check, lib
    check(lib().splash_readout_noalloc(s._ctx, buf.ptr, None))
    s.synchronize()
    return buf.read()


def test_readout_noalloc_equals_readout():
    import splash_amd as S
    cases = [("hex", 6, "HLLCplus", S.MIXED, 30), ("quad", 4, "HLLC", S.FP64, 12), ("hex", 4, "HLLE", S.MIXED, 7)]
    buf = None
    try:
        for kind, level, solver, prec, steps in cases:
            mesh = {"hex": S.icosphere_hex, "quad": S.quad_sphere}[kind](level)
            cls = getattr(S, "MUSCL_" + solver)
            if solver == "HLLE":
                U0 = S.ic.isothermal_state(mesh.face_centres(), seed=4)
                s = cls(mesh, U0, 4, GAM, precision=prec)
            else:
                U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=4, omega=(0.1, 0, 0.3))
                s = cls(mesh, U0, 5, GAM, 0.0, precision=prec)
            s.do_steps(1e-4, steps)
            if buf is None:
                buf = _DeviceBuffer(9)
            got = _noalloc(S, s, buf)
            want = np.concatenate([s.diagnostics(), s.get_light_curves() if solver != "HLLE" else np.zeros(4)])
            print(f"{kind}({level}) {solver}: {got}")
            assert np.array_equal(got, want)
            # a second call on the same state gives the same values
            assert np.array_equal(_noalloc(S, s, buf), want)
            del s
    finally:
        if buf is not None:
            buf.free()


def test_readout_noalloc_rejects_null_output():
    import splash_amd as S
    from splash_amd._lib import lib
    mesh = S.icosphere_hex(2)
    s = S.MUSCL_HLLCplus(mesh, S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=1), 5, GAM, 0.0)
    assert lib().splash_readout_noalloc(s._ctx, None, None) == -1
