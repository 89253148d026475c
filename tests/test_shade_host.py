"""The render kernels' camera and framebuffer helpers, built for the host (no GPU).

tests/cpp/shade_probe.cpp includes dr_shade.h as the kernels do and checks, on the CPU:
camera rays from the per-request constants (frame_k + camera_ray) bit for bit against the
oracle's camera ray (DumbRenderer.cu:190-232 with each lens) for every lens and boxing mode;
the reciprocal division (udiv) the kernels use for uniform divisors against integer division;
block_pixel against the 8 x 16 block layout (BlockBasedFrameBuffer.impl.cuh:125-143)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def probe(tmp_path_factory, oracle):
    if not os.path.exists(HIPCC) or shutil.which("g++") is None:
        pytest.skip("hipcc not available")
    oracle_dir = os.path.join(ROOT, "oracle")
    exe = str(tmp_path_factory.mktemp("shade_probe") / "shade_probe")
    subprocess.run([HIPCC, "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-x", "c++", os.path.join(HERE, "cpp", "shade_probe.cpp"), "-x", "none", "-L" + oracle_dir,
                    "-loracle", "-Wl,-rpath," + oracle_dir, "-o", exe], check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("mode", ["udiv", "camera", "blocks"])
def test_shade_helpers_on_host(probe, mode):
    r = subprocess.run([probe, mode], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
