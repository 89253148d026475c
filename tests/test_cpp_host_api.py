"""The C++ host mirror (include/dumbray.hpp: Octree, BlockBasedFrameBuffer<8,16>,
DumbRenderer::renderBlocksGPU) builds with hipcc (CPU test) and, on the GPU,
renders exactly what the Python mirror renders over the same C ABI."""
import os
import subprocess

import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "host_api_demo")


def test_cpp_host_api_builds():
    assert os.path.exists(EXE), "run __graft_entry__.build() first"


def _write_scene(path, sc, W, H, spp, depth, cull, seed):
    mats = D.plugins.material_table(sc.materials)[: len(sc.materials)]     # dr_material rows (48 B)
    lights = D.plugins.light_table(sc.lights)[: len(sc.lights)]           # dr_light rows (64 B)
    hdr = np.array([len(sc.tris), len(sc.materials), len(sc.lights), W, H, spp[0], spp[1], depth, int(cull),
                    (16 << 8) | 12], np.int32)
    with open(path, "wb") as f:
        for a in (hdr, sc.tris.astype(np.float32), sc.mat_id.astype(np.int32), mats, lights,
                  sc.camera.as_array()[:13], np.array([seed], np.uint64)):
            f.write(np.ascontiguousarray(a).tobytes())


@pytest.mark.gpu
def test_cpp_mirror_renders_like_python(dev, tmp_path):
    import torch
    sc = S.icosphere_scene(level=3, seed=3)
    W, H, spp, depth, seed = 50, 34, (2, 2), 5, 77
    scene_bin, out_bin = str(tmp_path / "scene.bin"), str(tmp_path / "out.bin")
    _write_scene(scene_bin, sc, W, H, spp, depth, True, seed)
    print("[case] C++ host mirror 50x34", flush=True)
    r = subprocess.run([EXE, scene_bin, out_bin], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "batch equal" in r.stdout        # RenderBatch (C++ mirror) = renderBlocksGPU, checked in the demo
    raw = np.fromfile(out_bin, dtype=np.uint8)
    nodes = int(raw[:4].view(np.int32)[0])
    img_cpp = raw[4:].view(np.float32).reshape(H, W, 4)
    t = D.Octree(16, 12)
    t.push(sc.tris)
    t.build()
    assert nodes == t.getNodeCount()
    g = D.DumbRenderer(t.upload(), sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, spp, depth, seed=seed)
    f = g.render()
    torch.cuda.synchronize()
    img_py = f.to_image().cpu().numpy()
    assert img_py[..., :3].max() > 0
    assert np.array_equal(img_cpp.view(np.int32), img_py.view(np.int32))
