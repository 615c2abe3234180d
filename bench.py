"""Throughput of the SPLASH MUSCL step on one MI355X.

Default workload (C3, the mixed-precision configuration): quad_sphere(9)
(1,572,864 faces, 6,291,456 half-edges), dim 5, MUSCL_HLLC, gam3d = 4/3
(gam = 2 - 1/gam3d), omega_ns = 0.1, accretion on (rate 1e-3, e_acc 0.1,
omega_acc 0.8, tilt 0.3 rad, width 0.1), friction on with alpha = 1e-6 (the
shipped parameter file's values), dt0 = 1e-4 (the CFL clamp may lower it:
the effective dt is reported), 1000 steps; seeded state Sigma = 1 +
0.1 N(0,1), Pi = 0.1 Sigma, solid-body rotation Omega = 0.3 about z plus
0.01 N(0,1) tangential noise (at Omega = 0.3 the effective gravity stays
positive, so the friction dt limit is live and stop_check stays clear; at
Omega = 0.5, `--state rotation05`, the reference driver would stop on step
1).  `--workload` selects the other single-mesh configurations (c2:
icosphere_hex(7) HLLC+ at Omega = 0.5; c5a / c5b: uv_sphere(2048,1024) /
icosphere_hex(9) HLLC+ with a spherical-harmonic density).

Both precision paths run the same steps from the same state.  Each timed
region (exactly --steps steps) follows --warmup untimed steps of the measured
solver and about 40 ms of steps of a separate pacer solver with the same
kernels, so the GPU starts it at the working clock of this kernel mix:
  value       = face-updates/s of the FP32-flux / FP64-state path,
  fp64        = face-updates/s of the plain FP64 path,
  l1_err      = relative L1 density error of the FP32-flux path vs FP64,
  driver_loop = ms per iteration of the reference driver's per-step loop
                (do_step, the diagnostics line's sums, time(),
                get_light_curves()) on the FP32-flux path, after the timed
                region.
The roofline object covers the dominant kernel of the FP32-flux path, timed
with HIP events on the run's stream, against the 8 TB/s peak and against a
device-to-device copy timed on the same GPU (copy_GBs); the CPU baseline is
the oracle (the C restatement of the reference step) on a bounded sample of
the same workload.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "muscl-scheme-on-spherical-mesh-wip_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
KERNELS = ["finalize_dt", "stage_1", "stage_2"]


def kernel_ms(prof, mixed):
    """Per-launch mean ms of one step's kernels and of the whole step from
    profile_step; the FP32-flux path launches no dt kernel (its stage 1 forms
    dt), so it lists the two stage kernels only."""
    names = KERNELS[1:] if mixed else KERNELS
    out = {k: float(prof[KERNELS.index(k)]) for k in names}
    out["step"] = float(prof[3])
    return out

from workloads import ACCRETION, GAM, STATES, WORKLOADS, c4_runs, initial_state, make_mesh  # noqa: E402

# HBM bytes per launch of the FP32-flux stage kernels, from rocprofv3 --pmc
# FETCH_SIZE and --pmc WRITE_SIZE runs of this script, FETCH_SIZE scaled by
# 2.0 and WRITE_SIZE by 1.0 (a 256 MiB device copy reads 128 MiB of
# FETCH_SIZE and 256 MiB of WRITE_SIZE).  Each entry holds for the kernel
# sources whose hash it names (kernel_source_hash); for any other source tree
# the bench line reports traffic as null.  The summary with the commands is
# under profiles/.
MEASURED_TRAFFIC = {
    ("c3", "rotation03", "stage_1", "9fe017f0c5be1ffd"): 432_350_618,
    ("c3", "rotation03", "stage_2", "9fe017f0c5be1ffd"): 603_728_486,
    ("c5b", "harmonics", "stage_1", "9fe017f0c5be1ffd"): 1_008_276_070,
    ("c5b", "harmonics", "stage_2", "9fe017f0c5be1ffd"): 1_281_739_571,
    ("c5a", "harmonics", "stage_1", "9fe017f0c5be1ffd"): 586_104_525,
    ("c5a", "harmonics", "stage_2", "9fe017f0c5be1ffd"): 798_778_778,
    ("c2", "rotation05", "stage_1", "9fe017f0c5be1ffd"): 65_029_837,   # footprint fits the MALL:
    ("c2", "rotation05", "stage_2", "9fe017f0c5be1ffd"): 81_832_448,   # L2-to-fabric bytes, not HBM
}
KERNEL_SOURCES = ["stage_mixed.hip", "stage_kernels.inc", "engine_types.hpp", "launch.hpp", "layout.cpp",
                  "run_setup.cpp"]


def kernel_source_hash():
    """First 16 hex digits of the SHA-256 of the sources that decide what the
    mixed stage kernels read and write."""
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "muscl-scheme-on-spherical-mesh-wip_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def algorithmic_bytes(kernel, F, H, d=5):
    """Canonical algorithmic bytes of one launch: FP64 state and geometry,
    every array read once and written once, no intermediates.  Per half-edge
    104 B (stencil faces 16, betas 32, H+- 16, BM 8, neighbour 4, half of the
    56 B per-edge normal/midpoint/arc), per face 36 B (area, unit centre,
    offset) and one state pass is 8*d B per face: stage 1 reads U0 and writes
    U_half, stage 2 reads U_half and U0 and writes U1.  The two stages sum to
    2*(104*H + 36*F) + 40*d*F per step."""
    if kernel == "stage_1":
        return 104 * H + 36 * F + 2 * 8 * d * F
    if kernel == "stage_2":
        return 104 * H + 36 * F + 3 * 8 * d * F
    return 0


NO_GRAPH = False  # --no-graph: every step launched on its own (kernel traces)


def params(S, accretion, **extra):
    return S.default_params(setup_threads=16, no_graph=int(NO_GRAPH), **(ACCRETION if accretion else {}), **extra)


def run_path(S, mesh, U0, wl, precision, steps, warmup, dt0):
    kind, margs, solver, omega_ns, acc, ic, _, _ = wl

    def make():
        return getattr(S, "MUSCL_" + solver)(mesh, U0, 5, GAM, omega_ns, acc, params=params(S, acc),
                                             precision=precision)

    s = make()
    pacer = make()  # a separate solver: its steps warm the GPU, s is not touched by them
    # the first do_steps call captures the step graph after its own steps,
    # so the warm-up (at least one step) keeps the capture out of the timing
    s.do_steps(dt0, warmup)
    s.synchronize()
    pace(pacer, dt0)
    t0 = time.perf_counter()
    s.do_steps(dt0, steps)
    s.synchronize()
    el = time.perf_counter() - t0
    del pacer  # after the timed region: its teardown waits on the GPU
    return s, el


DEVICE_WARMUP = "right before each timed region, ~40 ms of steps of a separate solver object with the same " \
                "kernels (same mesh, initial state and precision; for the ensemble, one HLLC+ run on its " \
                "largest mesh); the measured object's state and step count are untouched. After an idle gap " \
                "the GPU needs ~20 ms of load to reach its working clock; a streaming-copy warm-up left the " \
                "first timed steps ~7% slower than sustained, this one matches the sustained rate"


def pace(pacer, dt0, ms=40.0):
    """Step the separate solver (or ensemble) `pacer` for about `ms`
    milliseconds, waiting on it every 16 steps, so the GPU enters the
    timed region that follows at the working clock of this kernel mix."""
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < ms:
        pacer.do_steps(dt0, 16)
        pacer.synchronize()


def copy_bandwidth(S, reps=20):
    """The library's streaming copy kernel on 2 x 1 GiB, `reps` times: its
    bandwidth in GB/s (the achievable-bandwidth ceiling beside the roofline)."""
    import ctypes as C
    from splash_amd._lib import check, lib
    gbs = C.c_double()
    check(lib().splash_copy_bandwidth(1 << 30, reps, C.byref(gbs)))
    return gbs.value


def profile(s, dt0, n):
    """Per-kernel mean durations over n steps, timed with HIP events bound to
    each dispatch.  64 graph-replayed steps are queued first, with no host
    wait before the timed launches, so the GPU is at its working clock (after
    an idle gap it needs several milliseconds to return to it)."""
    s.do_steps(dt0, 64)
    return s.profile_step(dt0, n)


def driver_loop(s, dt0, n):
    """The reference driver's per-step pattern (test.cpp's loop): do_step,
    then the diagnostics line's sums, time() and get_light_curves(), one
    eager step at a time.  Returns (ms per loop iteration, ms per do_step
    alone), each over n iterations after a short warm-up."""
    def one(readout):
        s.do_step(dt0)
        if readout:
            s.diagnostics()
            s.time()
            s.get_light_curves()

    out = []
    for readout in (True, False):
        for _ in range(10):
            one(readout)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            one(readout)
        s.synchronize()
        out.append((time.perf_counter() - t0) / n * 1e3)
    return out


def ensemble_bench(S, steps, warmup, dt0, profile_steps):
    """The C4 ensemble: one launch per stage for all 64 runs, its patch shape
    chosen by the library's tuner (patch_faces = -1), against the same runs
    advanced one after another by single solvers (FP32-flux and FP64, default
    patch shape; results do not depend on it)."""
    runs = c4_runs(S)
    p = params(S, False)
    t_setup = time.perf_counter()
    ens = S.Ensemble([(m, U0, GAM, 0.0, params(S, False, patch_faces=-1)) for m, U0, _ in runs], 5, S.HLLCPLUS)
    t_setup = time.perf_counter() - t_setup
    n_runs, F, H, blocks = ens.info()
    tuned_faces, tuned_threads = ens.patch_shape()
    # pacers: single HLLC+ runs on the largest mesh, one per precision (the
    # mixed one runs the ensemble's stage kernels)
    big = max(range(n_runs), key=lambda r: runs[r][0].n_faces)
    pacers = {prec: S.MUSCL_HLLCplus(runs[big][0], runs[big][1], 5, GAM, 0.0, False, params=p, precision=prec)
              for prec in (S.MIXED, S.FP64)}
    ens.do_steps(dt0, warmup)
    ens.synchronize()
    pace(pacers[S.MIXED], dt0)
    t0 = time.perf_counter()
    ens.do_steps(dt0, steps)
    ens.synchronize()
    el = time.perf_counter() - t0
    Ue = [ens.get_state(r) for r in range(n_runs)]
    stops = sum(ens.get_stop_check(r) for r in range(n_runs))
    ens.do_steps(dt0, 64)
    prof = ens.profile_step(dt0, profile_steps)
    del ens
    # the same runs one at a time
    seq = {}
    num = den = 0.0
    linf = 0.0
    for prec, name in ((S.MIXED, "mixed"), (S.FP64, "fp64")):
        tot = 0.0
        for r, (m, U0, _) in enumerate(runs):
            s = S.MUSCL_HLLCplus(m, U0, 5, GAM, 0.0, False, params=p, precision=prec)
            s.do_steps(dt0, warmup)
            s.synchronize()
            pace(pacers[prec], dt0)
            t0 = time.perf_counter()
            s.do_steps(dt0, steps)
            s.synchronize()
            tot += time.perf_counter() - t0
            if prec == S.FP64:
                A = s.get_state()
                num += float(np.sum(np.abs(A[:, 0] - Ue[r][:, 0])))
                den += float(np.sum(np.abs(A[:, 0])))
                linf = max(linf, float(np.max(np.abs(A[:, 0] - Ue[r][:, 0])) / np.max(np.abs(A[:, 0]))))
            del s
        seq[name] = F * steps / tot
    del pacers
    ab = algorithmic_bytes("stage_2", F, H)
    return {
        "workload": f"c4: {n_runs} icosphere_hex runs, levels {sorted(set(lv for *_, lv in runs))} drawn with "
                    f"seed 4, F={F} H={H}, dim 5 MUSCL_HLLCplus, per-run seeds 4000+i, random axis, "
                    f"Omega~U(0,1), dt0={dt0}, {steps} steps",
        "value": F * steps / el, "unit": "face-updates/s", "ms_per_step": el / steps * 1e3,
        "one_run_at_a_time": {"mixed": seq["mixed"], "fp64": seq["fp64"]},
        "speedup_vs_fp64_runs": (F * steps / el) / seq["fp64"],
        "l1_err_vs_fp64": num / den, "linf_err_vs_fp64": linf,
        "kernel_ms": kernel_ms(prof, True),
        "stage_2_hbm_frac": ab / (prof[2] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "workgroups_per_stage": blocks, "runs_with_stop_check": int(stops),
        "patch_faces_tuned": tuned_faces, "block_threads_tuned": tuned_threads, "setup_s_incl_tuning": t_setup,
    }


def cpu_baseline(S, wl, U0, dt0, steps, threads):
    import oracle_lib as O
    kind, margs, solver, omega_ns, acc, ic, _, _ = wl
    m = O.Mesh(kind, *margs)
    p = O.make_params(**{k: v for k, v in ACCRETION.items()}) if acc else O.make_params()
    osolver = {"HLLCplus": O.HLLCPLUS, "HLLC": O.HLLC}[solver]
    ref = O.Solver(m, U0, 5, osolver, GAM, omega_ns, p, threads=threads)
    ref.do_step(dt0)  # warm-up, as in the reference timing protocol
    t0 = time.perf_counter()
    for _ in range(steps):
        ref.do_step(dt0)
    el = time.perf_counter() - t0
    return m.nf * steps / el, el


def memcpy_bandwidth(mib=1024, reps=10):
    """HBM bandwidth of the runtime's own copy on this GPU: a device-to-device hipMemcpy of
    `mib` MiB (read + write bytes per second), timed with HIP events over
    `reps` copies after one warm-up copy, through the HIP runtime the solver
    library itself loaded."""
    import ctypes as C
    with open("/proc/self/maps") as maps:  # the runtime libsplash_amd.so is linked against
        path = next((ln.split()[-1] for ln in maps if "libamdhip64" in ln), "libamdhip64.so.7")
    hip = C.CDLL(path)
    nbytes = mib * (1 << 20)

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    src, dst = C.c_void_p(), C.c_void_p()
    ok(hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)), "hipMalloc")
    ok(hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)), "hipMalloc")
    try:
        ok(hip.hipMemset(src, 0, C.c_size_t(nbytes)), "hipMemset")
        ok(hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3), "hipMemcpy")  # 3 = device to device
        e0, e1 = C.c_void_p(), C.c_void_p()
        ok(hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
        ok(hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
        ok(hip.hipEventRecord(e0, None), "hipEventRecord")
        for _ in range(reps):
            ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), 3, None), "hipMemcpyAsync")
        ok(hip.hipEventRecord(e1, None), "hipEventRecord")
        ok(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float()
        ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    return 2 * nbytes * reps / (ms.value * 1e-3) / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(k for k in WORKLOADS if k != "c1"), default="c3")
    ap.add_argument("--state", choices=["rotation03", "rotation05", "harmonics"], default=None,
                    help="initial state (default: the workload's own)")
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: the workload's step count)")
    ap.add_argument("--warmup", type=int, default=20, help="untimed steps before the timed ones (at least 1)")
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--cpu-face-updates", type=float, default=5e7,
                    help="size of the CPU baseline sample in face-updates (whole steps, at least one)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-run ensemble (C4) measurement")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--driver-steps", type=int, default=200,
                    help="iterations of the reference driver's per-step loop (do_step + readouts) to time")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step on its own instead of replaying captured HIP graphs")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    args.warmup = max(1, args.warmup)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration_mib:
        import torch
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy calibration: {args.copy_calibration_mib} MiB read + {args.copy_calibration_mib} MiB written",
              file=sys.stderr)

    global NO_GRAPH
    NO_GRAPH = args.no_graph
    import splash_amd as S
    wl = WORKLOADS[args.workload]
    if args.state:
        wl = wl[:5] + (args.state,) + wl[6:]
    kind, margs, solver, omega_ns, acc, ic, wl_steps, _ = wl
    steps = args.steps or wl_steps
    mesh = make_mesh(S, kind, margs)
    F, H = mesh.n_faces, mesh.n_halfedges
    U0 = initial_state(S, mesh, ic)

    s64, el64 = run_path(S, mesh, U0, wl, S.FP64, steps, args.warmup, args.dt)
    U64 = s64.get_state()
    t64 = s64.time()
    prof64 = profile(s64, args.dt, args.profile_steps)
    del s64
    smx, elmx = run_path(S, mesh, U0, wl, S.MIXED, steps, args.warmup, args.dt)
    Umx = smx.get_state()
    stop = smx.get_stop_check()
    tmx = smx.time()
    dt_eff = smx.last_dt()
    patch_faces, block_threads = smx.patch_shape()
    prof = profile(smx, args.dt, args.profile_steps)
    loop_ms, step_ms = driver_loop(smx, args.dt, args.driver_steps)
    del smx

    l1 = float(np.sum(np.abs(Umx[:, 0] - U64[:, 0])) / np.sum(np.abs(U64[:, 0])))
    linf = float(np.max(np.abs(Umx[:, 0] - U64[:, 0])) / np.max(np.abs(U64[:, 0])))
    value = F * steps / elmx
    value64 = F * steps / el64

    k = int(np.argmax(prof[:3]))
    name = KERNELS[k]
    ab = algorithmic_bytes(name, F, H)
    achieved = ab / (prof[k] * 1e-3) / 1e9
    copy_gbs = copy_bandwidth(S)  # the library's streaming copy kernels, the faster form
    memcpy_gbs = memcpy_bandwidth()
    step_bytes = 2 * (104 * H + 36 * F) + 40 * 5 * F
    src_hash = kernel_source_hash()
    traffic = MEASURED_TRAFFIC.get((args.workload, ic, name, src_hash))

    mesh_name = {"hex": "icosphere_hex", "quad": "quad_sphere", "uv": "uv_sphere"}[kind]
    state = STATES[ic]
    sources = ("accretion on (rate 1e-3, e_acc 0.1, omega_acc 0.8, tilt 0.3, width 0.1), friction alpha 1e-6"
               if acc else "accretion off")
    line = {
        "metric": "face-updates/s, FP32-flux vs FP64 GPU at <=1e-5 L1 error; %HBM roofline",
        "value": value,
        "unit": "face-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elmx / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-flux/f64-state",
        "data": "synthetic (seeded splitmix64/Box-Muller state on the generated mesh)",
        "config": {
            "workload": f"{args.workload}: {mesh_name}{margs} F={F} H={H}, dim 5 MUSCL_{solver}, gam3d 4/3, "
                        f"omega_ns={omega_ns}, {state}, {sources}, dt0={args.dt}",
            "faces": F, "halfedges": H, "dt0": args.dt, "dt_effective": dt_eff,
            "patch_faces": patch_faces, "block_threads": block_threads,
        },
        "l1_err_vs_fp64": l1,
        "linf_err_vs_fp64": linf,
        "speedup_vs_fp64": value / value64,
        "fp64": {"value": value64, "ms_per_step": el64 / steps * 1e3,
                 "kernel_ms": kernel_ms(prof64, False),
                 "step_hbm_frac": step_bytes / (prof64[3] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "kernel_ms": kernel_ms(prof, True),
        "step_algorithmic_bytes": step_bytes,
        "step_achieved_GBs": step_bytes / (prof[3] * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_source_hash": src_hash,
                     "algorithmic_bytes": ab, "copy_GBs": copy_gbs, "memcpy_GBs": memcpy_gbs,
                     "frac_of_copy": achieved / copy_gbs,
                     "traffic_GBs": traffic / (prof[k] * 1e-3) / 1e9 if traffic else None},
        "stop_check": bool(stop),
        "device_warmup": DEVICE_WARMUP,
        "sim_time": [t64, tmx],
        "driver_loop": {"steps": args.driver_steps, "ms_per_step": loop_ms, "do_step_ms": step_ms,
                        "readout_ms": loop_ms - step_ms,
                        "loop": "FP32-flux do_step, then diagnostics sums, time() and get_light_curves(), "
                                "eager, after the timed region"},
    }
    if not args.no_ensemble:
        line["ensemble"] = ensemble_bench(S, args.ensemble_steps, args.warmup, args.dt, args.profile_steps)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu_steps = max(1, int(round(args.cpu_face_updates / F)))
        cpu_v, cpu_el = cpu_baseline(S, wl, U0, args.dt, cpu_steps, threads)
        line["cpu_baseline"] = {"value": cpu_v, "unit": "face-updates/s", "cores": threads, "kind": "port",
                                "sample": f"{cpu_steps} FP64 steps of the same workload after 1 warm-up "
                                          f"step, oracle (C restatement) with {threads} OpenMP threads, "
                                          f"{cpu_el:.1f} s"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
