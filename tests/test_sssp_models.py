"""Host models of the SSSP bucket loops (CPU).

Both delta-stepping engines drive their phases with a state machine that
runs on the device between phases (graphbench_amd/csrc/gb_sssp_ctl.hpp).
The models in tests/models compile that header for the host, emulate each
action (light/heavy expansion, far scan, mask clearing) sequentially on
random graphs -- hubs split into chunks, duplicate sources, bucket widths
from 1 to beyond the largest weight -- and compare every distance with
Dijkstra.  They pin the loop logic (termination, counter resets, queue
double-buffering, bucket skipping) independently of the GPU.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "graphbench_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.parametrize("model,trials", [("sssp_ctl_model", 300), ("msssp_ctl_model", 300)])
def test_state_machine_model_matches_dijkstra(tmp_path, model, trials):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = tmp_path / model
    src = os.path.join(HERE, "models", model + ".hip")
    # the models run on the host; the device flags match the library's build
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang",
                    "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", CSRC, "-o", str(exe),
                    src], check=True, capture_output=True)
    r = subprocess.run([str(exe), str(trials)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok"), r.stdout
