"""Progressive rendering and image output on the GPU.

- dr_frame_to_rgba8 (device kernel) against the oracle's byte conversion, exactly, on
  crafted frames (special values, sizes that are not whole blocks);
- Images.saveBufferPNG read back equals the device bytes;
- BufferedRenderProcess (Python mirror) resizes the buffer, runs a target number of
  iterations of the reference's example scene and ends on the oracle's 1/k blend of
  those iterations; double buffers hold one iteration each;
- the C++ host mirror's workflow (DumbRenderContext::buildFromFile, BufferedRenderProcess,
  Images::saveBufferPNG) writes the same PNG bytes the Python mirror produces."""
import os
import subprocess
import threading

import numpy as np
import pytest

import dumbray_amd as D
import oracle
from oracle import scene_file as SF

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(ROOT, "tests", "golden", "example_scene_0", "0.dumb")
DEMO = os.path.join(ROOT, "tests", "cpp", "scene_file_demo")
RTOL, ATOL = 1e-4, 1e-6


def _close(a, b):
    err = np.abs(a - b)
    bad = err > RTOL * np.abs(b) + ATOL
    assert not bad.any(), (int(bad.sum()), float(err.max()), np.argwhere(bad)[:5].tolist())


@pytest.mark.parametrize("w,h", [(8, 16), (13, 21), (100, 37)])
def test_frame_to_rgba8_matches_oracle(dev, w, h):
    import torch

    print(f"[case] frame_to_rgba8 {w}x{h}", flush=True)
    rng = np.random.default_rng(w + h)
    img = rng.uniform(-0.5, 1.5, (h, w, 4)).astype(np.float32)
    img.reshape(-1)[:: 7] = np.nan
    img.reshape(-1)[3:: 11] = 1.0 / 255.0
    img.reshape(-1)[5:: 13] = 0.5
    f = D.BlockBasedFrameBuffer(w, h)
    f.data.copy_(f.from_image(torch.from_numpy(img).cuda()))
    got = D.images.frame_to_rgba8(f)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy(), oracle.frame_bytes(img))


def test_save_png_matches_device_bytes(dev, tmp_path):
    import torch
    from PIL import Image

    sc_ctx = D.DumbRenderContext.fromFile(EXAMPLE)
    g = sc_ctx.renderer(40, 24)
    f = g.render()
    torch.cuda.synchronize()
    path = str(tmp_path / "f.png")
    assert D.images.saveBufferPNG(f, path) == D.images.IMAGES_NO_ERROR
    np.testing.assert_array_equal(np.asarray(Image.open(path)), D.images.frame_to_rgba8(f).cpu().numpy())


def _oracle_renderer(W, H):
    sf = SF.load(EXAMPLE)
    return oracle.Renderer(oracle.Octree(sf.tris, 8, 24), sf.mat_id, sf.materials, sf.lights, sf.cam, W, H, sf.fsaa,
                           sf.max_bounces, ignore_backfaces=sf.ignore_backfaces, normals=sf.normals, boxing=sf.boxing)


def _run(process, target, timeout=60.0):
    done = threading.Event()
    counts = {"iter": 0, "done": 0, "error": 0}

    def on_iter(_):
        counts["iter"] += 1

    def on_done(_):
        counts["done"] += 1
        done.set()

    def on_error(_):
        counts["error"] += 1

    process.setIterationCompletionCallback(on_iter)
    process.setRenderCompletionCallback(on_done)
    process.setErrorOnIteration(on_error)
    process.setTargetIterations(target)
    process.start()
    assert done.wait(timeout), "render did not complete"
    process.end()
    return counts


def test_buffered_render_process_progressive(dev):
    import torch

    W, H = 64, 36
    ctx = D.DumbRenderContext.fromFile(EXAMPLE)
    g = ctx.renderer(W, H)
    frame = D.BlockBasedFrameBuffer(8, 8)
    p = D.BufferedRenderProcess()
    p.setBuffer(frame)
    p.setRenderer(g)
    p.setTargetResolution(W, H)
    print("[case] BufferedRenderProcess 3 iterations 64x36", flush=True)
    counts = _run(p, 3)
    assert frame.getSize() == (W, H) and g.iteration() == 3
    assert counts == {"iter": 3, "done": 1, "error": 0}
    o = _oracle_renderer(W, H)
    acc = None
    for k in (1, 2, 3):
        acc = oracle.blend(acc, o.render_ref(iteration=k, nthreads=8), k)
    torch.cuda.synchronize()
    _close(frame.to_image()[..., :3].cpu().numpy(), acc)
    # raising the target continues the same progression
    print("[case] BufferedRenderProcess continue to 5", flush=True)
    counts = _run(p, 5)
    assert g.iteration() == 5 and counts["iter"] == 2
    for k in (4, 5):
        acc = oracle.blend(acc, o.render_ref(iteration=k, nthreads=8), k)
    _close(frame.to_image()[..., :3].cpu().numpy(), acc)


def test_buffered_render_process_double_buffers(dev):
    import time

    W, H = 48, 32
    ctx = D.DumbRenderContext.fromFile(EXAMPLE)
    g = ctx.renderer(W, H)
    a, b = D.BlockBasedFrameBuffer(W, H), D.BlockBasedFrameBuffer(W, H)
    p = D.BufferedRenderProcess()
    p.setDoubleBuffers(a, b)
    p.setRenderer(g)
    seen = threading.Event()
    n = {"iter": 0}

    def on_iter(_):
        n["iter"] += 1
        if n["iter"] >= 4:
            seen.set()

    p.setIterationCompletionCallback(on_iter)
    print("[case] BufferedRenderProcess double buffers", flush=True)
    p.start()
    assert seen.wait(60.0)
    p.end()
    time.sleep(0)
    ref = _oracle_renderer(W, H).render_ref(iteration=1, nthreads=8)
    for f in (a, b):                                  # each buffer holds one iteration-1 render
        _close(f.to_image()[..., :3].cpu().numpy(), ref)


def _lens_scene(tmp_path):
    """The example scene seen through the spherical segment lens."""
    import re
    text = open(EXAMPLE).read().replace('"files/cargo.obj"', '"%s"' % os.path.join(os.path.dirname(EXAMPLE), "files",
                                                                                     "cargo.obj"))
    text = re.sub(r'"lense"\s*:\s*\{[^}]*\}', '"lense": {"type": "spherical_segment", "angle": 40, "distance": 140, '
                  '"sensor": 2.5, "filter": [1.0, 0.9, 0.7]}', text)
    path = tmp_path / "lens.dumb"
    path.write_text(text)
    return str(path)


@pytest.mark.parametrize("scene", ["example", "textured", "spherical_segment_lens"])
def test_cpp_scene_file_workflow_matches_python(dev, tmp_path, scene):
    """The C++ mirror's workflow writes the Python mirror's PNG for the example scene, a
    textured scene file and the example scene through the spherical segment lens."""
    import torch
    from PIL import Image
    from test_scene_file import write_textured_scene

    assert os.path.exists(DEMO), "run __graft_entry__.build() first"
    path = {"example": lambda: EXAMPLE, "textured": lambda: write_textured_scene(tmp_path),
            "spherical_segment_lens": lambda: _lens_scene(tmp_path)}[scene]()
    W, H, N = 64, 36, 3
    out = str(tmp_path / "cpp.png")
    print(f"[case] C++ scene_file_demo {scene} 64x36 3 iterations", flush=True)
    r = subprocess.run([DEMO, path, str(W), str(H), str(N), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    if scene == "example":
        assert "triangles 1670" in r.stdout
    ctx = D.DumbRenderContext.fromFile(path)
    g = ctx.renderer(W, H)
    f = None
    for k in range(1, N + 1):
        f = g.render(f, k)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(np.asarray(Image.open(out)), D.images.frame_to_rgba8(f).cpu().numpy())


def test_progress_checkpoint_resumes_bit_exactly(dev, tmp_path):
    """saveProgress after 3 iterations, loadProgress into a fresh renderer, 2 more iterations:
    the frame equals 5 uninterrupted iterations bit for bit (iterations are keyed by their
    number); a renderer with other settings refuses the file."""
    import torch
    ctx = D.DumbRenderContext.fromFile(EXAMPLE)
    tree = ctx.octree()
    W, H = 40, 24
    print(f"[case] progress checkpoint example scene {W}x{H}", flush=True)
    a = ctx.renderer(W, H, tree=tree, samples_per_pixel=(2, 1))
    fa = D.BlockBasedFrameBuffer(W, H)
    a.setFrameBuffer(fa)
    for _ in range(5):
        assert a.iterate()
    b = ctx.renderer(W, H, tree=tree, samples_per_pixel=(2, 1))
    b.setFrameBuffer(D.BlockBasedFrameBuffer(W, H))
    for _ in range(3):
        assert b.iterate()
    path = str(tmp_path / "progress.npz")
    b.saveProgress(path)
    c = ctx.renderer(W, H, tree=tree, samples_per_pixel=(2, 1))
    fc = c.loadProgress(path)
    assert c.iteration() == 3 and c.getFrameBuffer() is fc
    for _ in range(2):
        assert c.iterate()
    torch.cuda.synchronize()
    assert c.iteration() == 5
    assert torch.equal(fa.data.view(torch.int32), fc.data.view(torch.int32))
    other = ctx.renderer(W, H, tree=tree, samples_per_pixel=(1, 1))
    with pytest.raises(D.DumbRayError):
        other.loadProgress(path)
    # another resolution is refused too, and the renderer keeps its own size
    small = ctx.renderer(W // 2, H, tree=tree, samples_per_pixel=(2, 1))
    with pytest.raises(D.DumbRayError):
        small.loadProgress(path)
    assert (small.width, small.height) == (W // 2, H) and small.getFrameBuffer() is None
