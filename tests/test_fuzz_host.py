"""Mutation fuzzing of the host-side parsers (scene files, OBJ meshes, PNG) under
AddressSanitizer and UndefinedBehaviorSanitizer: tests/fuzz/fuzz_host.cpp is built with
g++ from the library's own sources and run on mutations of the reference's example scene,
a grouped scene with includes, a textured scene and PNGs of every colour type.  A
sanitizer report or a crash fails the test.  Host only (no GPU)."""
import os
import shutil
import subprocess
import textwrap
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "dumbray_amd", "csrc")
EXAMPLE_DIR = os.path.join(HERE, "golden", "example_scene_0")
FLAGS = ["-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-ffp-contract=off"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def fuzzer(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("fuzz_build")
    srcs = [os.path.join(HERE, "fuzz", "fuzz_host.cpp"), os.path.join(CSRC, "dr_scene_file.cpp"),
            os.path.join(CSRC, "dr_png.cpp")]

    def compile_one(src):
        obj = str(d / (os.path.basename(src) + ".o"))
        subprocess.run(["g++", *FLAGS, "-c", src, "-o", obj], check=True, capture_output=True)
        return obj
    with ThreadPoolExecutor(3) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = str(d / "fuzz_host")
    subprocess.run(["g++", *FLAGS, *objs, "-lz", "-o", exe], check=True, capture_output=True)
    return exe


def _run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, env=ENV, timeout=600)
    report = r.stdout + r.stderr
    assert r.returncode == 0 and "runtime error" not in report and "Sanitizer" not in report, report[-4000:]
    return report


def test_fuzz_scene_text(fuzzer, tmp_path):
    import test_scene_file_gpu as G
    for name, text in (("scene.dumb", G.SCENE), ("inc.dumb", G.INC), ("mesh.obj", G.MESH)):
        (tmp_path / name).write_text(textwrap.dedent(text))
    print(_run(fuzzer, "scene", 2000, tmp_path, tmp_path / "scene.dumb", tmp_path / "inc.dumb"))
    print(_run(fuzzer, "scene", 300, EXAMPLE_DIR, os.path.join(EXAMPLE_DIR, "0.dumb")))


def test_fuzz_textured_scene(fuzzer, tmp_path):
    from test_scene_file import write_textured_scene
    path = write_textured_scene(tmp_path)
    print(_run(fuzzer, "scene", 1500, tmp_path, path))


def test_fuzz_obj(fuzzer, tmp_path):
    import test_scene_file_gpu as G
    scene = tmp_path / "s.dumb"
    scene.write_text('{"objects": [{"mesh": {"obj": "mesh.obj"}, "material": "m"}], '
                     '"materials": [{"name": "m", "type": "simple_stochastic"}]}')
    seed = tmp_path / "seed.obj"
    seed.write_text(textwrap.dedent(G.MESH))
    print(_run(fuzzer, "obj", 2000, tmp_path, scene, seed))
    print(_run(fuzzer, "obj", 150, tmp_path, scene, os.path.join(EXAMPLE_DIR, "files", "cargo.obj")))


def test_fuzz_png(fuzzer, tmp_path):
    from PIL import Image
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, (9, 7, 4), dtype=np.uint8)
    seeds = []
    for mode in ("RGBA", "RGB", "L", "LA", "1"):
        p = tmp_path / f"{mode}.png"
        Image.fromarray(a, "RGBA").convert(mode).save(p)
        seeds.append(p)
    p = tmp_path / "P.png"
    Image.fromarray(a[..., :3].copy(), "RGB").quantize(5).save(p, transparency=1)
    seeds.append(p)
    # Adam7, written here (Pillow writes no interlaced PNGs): each pass's rows, filter 0
    from test_images import _png
    raw = b""
    for x0, y0, dx, dy in [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]:
        sub = a[y0::dy, x0::dx]
        if sub.size:
            raw += b"".join(b"\x00" + sub[y].tobytes() for y in range(sub.shape[0]))
    p = tmp_path / "adam7.png"
    p.write_bytes(_png(7, 9, 8, 6, raw, interlace=1))
    seeds.append(p)
    print(_run(fuzzer, "png", 10000, tmp_path, *seeds))
