"""Throughput of the SPLASH MUSCL step on one MI355X.

Workload (default): icosphere_hex(7) (163,842 faces, 983,040 half-edges),
dim 5, MUSCL_HLLCplus, gam3d = 4/3 (gam = 2 - 1/gam3d), omega_ns = 0,
accretion off, dt0 = 1e-4; seeded state Sigma = 1 + 0.1 N(0,1), Pi = 0.1
Sigma, solid-body rotation Omega = 0.5 about z plus 0.01 N(0,1) tangential
noise.  Both precision paths run the same steps from the same state:
  value       = face-updates/s of the FP32-flux / FP64-state path,
  fp64        = face-updates/s of the plain FP64 path,
  l1_err      = relative L1 density error of the FP32-flux path vs FP64.
The roofline object covers the dominant kernel of the FP32-flux path, timed
with HIP events on the run's stream; the CPU baseline is the oracle (the C
restatement of the reference step) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "muscl-scheme-on-spherical-mesh-wip_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
GAM = 2 - 1 / (4 / 3)
KERNELS = ["finalize_dt", "stage_1", "stage_2"]


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


def run_path(S, mesh, U0, precision, steps, warmup, dt0):
    s = S.MUSCL_HLLCplus(mesh, U0, 5, GAM, 0.0, False, params=S.default_params(setup_threads=16),
                         precision=precision)
    s.do_steps(dt0, warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.do_steps(dt0, steps)
    s.synchronize()
    el = time.perf_counter() - t0
    return s, el


def cpu_baseline(level, U0, dt0, steps, threads):
    import oracle_lib as O
    m = O.Mesh("hex", level)
    ref = O.Solver(m, U0, 5, O.HLLCPLUS, GAM, 0.0, threads=threads)
    ref.do_step(dt0)  # warm-up, as in the reference timing protocol
    t0 = time.perf_counter()
    for _ in range(steps):
        ref.do_step(dt0)
    el = time.perf_counter() - t0
    return m.nf * steps / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
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

    import splash_amd as S
    mesh = S.icosphere_hex(args.level)
    F, H = mesh.n_faces, mesh.n_halfedges
    U0 = S.ic.adiabatic_state(mesh.face_centres(), GAM, seed=2, sigma_amp=0.1, pi_frac=0.1,
                              omega=(0.0, 0.0, 0.5), noise=0.01)

    s64, el64 = run_path(S, mesh, U0, S.FP64, args.steps, args.warmup, args.dt)
    U64 = s64.get_state()
    t64 = s64.time()
    prof64 = s64.profile_step(args.dt, args.profile_steps)
    del s64
    smx, elmx = run_path(S, mesh, U0, S.MIXED, args.steps, args.warmup, args.dt)
    Umx = smx.get_state()
    stop = smx.get_stop_check()
    tmx = smx.time()
    prof = smx.profile_step(args.dt, args.profile_steps)
    del smx

    l1 = float(np.sum(np.abs(Umx[:, 0] - U64[:, 0])) / np.sum(np.abs(U64[:, 0])))
    linf = float(np.max(np.abs(Umx[:, 0] - U64[:, 0])) / np.max(np.abs(U64[:, 0])))
    value = F * args.steps / elmx
    value64 = F * args.steps / el64

    k = int(np.argmax(prof[:3]))
    name = KERNELS[k]
    ab = algorithmic_bytes(name, F, H)
    achieved = ab / (prof[k] * 1e-3) / 1e9
    step_bytes = 2 * (104 * H + 36 * F) + 40 * 5 * F

    line = {
        "metric": "face-updates/s, FP32-flux vs FP64 GPU at <=1e-5 L1 error; %HBM roofline",
        "value": value,
        "unit": "face-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elmx / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-flux/f64-state",
        "data": "synthetic (seeded splitmix64/Box-Muller state on the generated mesh)",
        "config": {
            "workload": f"icosphere_hex({args.level}) F={F} H={H}, dim 5 MUSCL_HLLCplus, gam3d 4/3, "
                        f"Sigma=1+0.1N, Pi=0.1Sigma, Omega=0.5 z + 0.01N, dt0={args.dt}, accretion off",
            "faces": F, "halfedges": H, "dt0": args.dt,
        },
        "l1_err_vs_fp64": l1,
        "linf_err_vs_fp64": linf,
        "fp64": {"value": value64, "ms_per_step": el64 / args.steps * 1e3,
                 "kernel_ms": dict(zip(KERNELS + ["step"], map(float, prof64))),
                 "step_hbm_frac": step_bytes / (prof64[3] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "kernel_ms": dict(zip(KERNELS + ["step"], map(float, prof))),
        "step_algorithmic_bytes": step_bytes,
        "step_achieved_GBs": step_bytes / (prof[3] * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": ab},
        "stop_check": bool(stop),
        "sim_time": [t64, tmx],
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu_v, cpu_el = cpu_baseline(args.level, U0, args.dt, args.cpu_steps, threads)
        line["cpu_baseline"] = {"value": cpu_v, "unit": "face-updates/s", "cores": threads, "kind": "port",
                                "sample": f"{args.cpu_steps} FP64 steps of the same workload after 1 warm-up "
                                          f"step, oracle (C restatement) with {threads} OpenMP threads, "
                                          f"{cpu_el:.1f} s"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
