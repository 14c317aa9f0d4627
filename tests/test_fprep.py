"""The closed-form repetition of FP32 additions the raycast uses to skip
empty space (vcl-volume-scanner_amd/csrc/vs_fprep.hpp) against the
step-by-step loop, bit for bit, on the host.

The march's steps through empty regions must end on the reference's FP32
position and distance (tsdf_renderer.cu:68-153 performs each addition), so
the skip is exact by construction; this checks it over seeded random runs
(halfway steps, binade edges, zero crossings, fixed points, the march's own
ranges), once as built and once with the reciprocal estimate pushed to the
top of the device reciprocal's error. A deliberately unsafe estimate (1%
high) must be caught, which shows the check can fail.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "vcl-volume-scanner_amd", "csrc")


def _run(tmp_path, scale, n):
    exe = tmp_path / f"fprep_{abs(hash(scale))}"
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{CSRC}", os.path.join(HERE, "fprep_check.cpp"),
           "-o", str(exe)]
    if scale is not None:
        cmd.insert(1, f"-DVS_FPREP_RCP_SCALE={scale}")
    subprocess.run(cmd, check=True)
    return subprocess.run([str(exe), str(n)], capture_output=True, text=True)


@pytest.mark.parametrize("scale", [None, "1.0000002384185791f"])
def test_repeat_add_is_exact(tmp_path, scale):
    r = _run(tmp_path, scale, 600000)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    runs, jumped = int(r.stdout.split()[-5]), int(r.stdout.split()[-3])
    assert runs == 600000 and jumped > runs // 2  # most runs take the closed form


def test_unsafe_estimate_is_caught(tmp_path):
    r = _run(tmp_path, "1.01f", 100000)
    assert r.returncode != 0 and "mismatch" in r.stdout
