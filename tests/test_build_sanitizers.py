"""The parallel host octree builder (dr_octree_build.hip, host code) under ThreadSanitizer
and under AddressSanitizer + UndefinedBehaviorSanitizer: 8 threads on a 120k-triangle soup
(nodes above the team-split size, split 8 / depth 24) and on a 256x256 heightfield
(split 16 / depth 12).  No sanitizer report, and the node count equals the library's own
build.  g++ compiles the builder's source as C++ (HIP headers for the host types only)."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "dumbray_amd", "csrc")
BASE = ["g++", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-ffp-contract=off", "-pthread",
        "-I/opt/rocm/include", "-I" + CSRC, "-D__HIP_PLATFORM_AMD__", "-x", "c++",
        os.path.join(CSRC, "dr_octree_build.hip"), "-x", "none", os.path.join(HERE, "cpp", "build_octree_main.cpp")]
SANITIZERS = {"tsan": ["-fsanitize=thread"], "asan_ubsan": ["-fsanitize=address,undefined"]}
CASES = [("soup120k", lambda: S.random_soup(120000, seed=3, extent=20.0, size=1.0), 8, 24),
         ("heightfield256", lambda: S.heightfield(256), 16, 12)]


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the HIP headers not available")
    d = tmp_path_factory.mktemp("build_sanitizers")

    def make(name):
        exe = str(d / name)
        subprocess.run(BASE + SANITIZERS[name] + ["-o", exe], check=True, capture_output=True)
        return name, exe
    with ThreadPoolExecutor(2) as ex:
        return dict(ex.map(make, SANITIZERS))


@pytest.mark.parametrize("san", list(SANITIZERS))
@pytest.mark.parametrize("name,make,split,depth", CASES, ids=[c[0] for c in CASES])
def test_builder_clean_under_sanitizer(builds, tmp_path, san, name, make, split, depth):
    tris = np.ascontiguousarray(make(), np.float32)
    path = tmp_path / "tris.bin"
    tris.tofile(path)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([builds[san], str(path), "8", str(split), str(depth)], capture_output=True, text=True,
                       env=env, timeout=600)
    report = r.stdout + r.stderr
    assert r.returncode == 0 and "Sanitizer" not in report and "runtime error" not in report, report[-4000:]
    t = D.Octree(split, depth)
    t.push(tris)
    t.build(nthreads=8)
    assert report.split()[:4] == ["rc", "0", "nodes", str(t.getNodeCount())], report
