"""bench.py's byte model, traffic records and command-line contract (no GPU).

- The canonical algorithmic bytes per face-update: per half-edge 104 B and per
  face 36 B for each of the two stages, plus five FP64 state passes
  (stage 1 reads U0 and writes U_half, stage 2 reads U_half and U0 and
  writes U1): 1520 B on hexagonal meshes (H = 6F) and 1104 B on quad meshes
  (H = 4F).
- Every measured-traffic record names the hash of the current kernel
  sources, so a change to the kernels that is not re-measured fails here
  instead of reporting stale traffic.
- More than one GPU is refused with a message and exit status 2.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from workloads import STATES, WORKLOADS  # noqa: E402


@pytest.mark.parametrize("valence, per_face_update", [(6, 1520), (4, 1104)])
def test_algorithmic_bytes_per_face_update(valence, per_face_update):
    F = 1000
    H = valence * F
    step = bench.algorithmic_bytes("stage_1", F, H) + bench.algorithmic_bytes("stage_2", F, H)
    assert step == per_face_update * F
    # the stage-2 launch carries one more state pass (reading U0) than stage 1
    assert bench.algorithmic_bytes("stage_2", F, H) - bench.algorithmic_bytes("stage_1", F, H) == 40 * F
    assert bench.algorithmic_bytes("finalize_dt", F, H) == 0


def test_traffic_records_name_the_current_kernel_sources():
    h = bench.kernel_source_hash()
    assert len(h) == 16
    assert bench.MEASURED_TRAFFIC, "no traffic recorded"
    for (workload, state, kernel, src_hash), nbytes in bench.MEASURED_TRAFFIC.items():
        assert src_hash == h, f"{workload}/{kernel} was measured on other kernel sources ({src_hash} != {h})"
        assert workload in WORKLOADS and state in STATES and kernel in ("stage_1", "stage_2")
        assert WORKLOADS[workload][5] == state, "recorded for the workload's own initial state"
        assert nbytes > 0
    # the default workload's dominant kernel has its traffic
    assert ("c3", WORKLOADS["c3"][5], "stage_2", h) in bench.MEASURED_TRAFFIC


def test_pacer_steps_only_the_pacer():
    class Pacer:
        steps = 0
        waits = 0

        def do_steps(self, dt0, n):
            assert dt0 == 1e-4 and n == 16
            self.steps += n

        def synchronize(self):
            self.waits += 1

    p = Pacer()
    bench.pace(p, 1e-4, ms=2.0)
    assert p.steps >= 16 and p.waits == p.steps // 16


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
