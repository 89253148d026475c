"""The failure path of the render operator.  The reference's renderBlocksGPU renders the
blocks on the CPU when the GPU path fails and returns false (DumbRenderer.cu:109-111), and
BlockRenderer::iterateGPU then stops using the device (BlockRenderer.cu:89-91).  Here a
refused call returns a status and a message, leaves the frame buffer and the gradient
buffers untouched, and leaves the library usable, so the caller can fall back the same way."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S
from dumbray_amd._lib import lib

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _renderer(W=24, H=20):
    sc = S.spheres_scene(8, seed=0)
    t = D.Octree(16, 12)
    t.push(sc.tris)
    t.build()
    return sc, D.DumbRenderer(t.upload(), sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, (2, 2), 3)


def test_refused_render_leaves_frame_untouched(dev):
    import torch
    sc, g = _renderer()
    frame = D.BlockBasedFrameBuffer(g.width, g.height)
    nb = frame.getBlockCount()
    cases = {
        "material table smaller than the ids in use": dict(nmat=1),
        "unknown lens": dict(lens=9),
        "unknown plugin set": dict(plugins=7),
        "max_bounces": dict(max_bounces=65),
        "nlights": dict(nlights=17),
        "rr_survive": dict(rr_survive=0.0),
    }
    s = torch.cuda.current_stream()
    for msg, change in cases.items():
        p = g.params(1, 0)
        for k, v in change.items():
            setattr(p, k, v)
        frame.data.fill_(7.0)
        rc = lib().dr_render_blocks(g.scene.handle, ctypes.byref(p), ctypes.c_void_p(frame.data.data_ptr()), 0, nb,
                                    ctypes.c_void_p(s.cuda_stream))
        torch.cuda.synchronize()
        assert rc == 1 and msg.encode() in lib().dr_last_error(), (msg, rc, lib().dr_last_error())
        assert bool((frame.data == 7.0).all()), msg
    p = g.params(1, 0)
    for lo, hi in ((0, nb + 1), (-1, 2), (3, 2)):
        rc = lib().dr_render_blocks(g.scene.handle, ctypes.byref(p), ctypes.c_void_p(frame.data.data_ptr()), lo, hi,
                                    ctypes.c_void_p(s.cuda_stream))
        assert rc == 1 and b"block range" in lib().dr_last_error()
    torch.cuda.synchronize()
    assert bool((frame.data == 7.0).all())
    # the adjoint of the reference plugin set is refused and adds nothing
    ga = torch.full((g.nmat, 3), 5.0, dtype=torch.float64, device="cuda")
    ge = torch.full((g.nlights, 3), 5.0, dtype=torch.float64, device="cuda")
    p.plugins = 1
    rc = lib().dr_backward(g.scene.handle, ctypes.byref(p), ctypes.c_void_p(frame.data.data_ptr()),
                           ctypes.c_void_p(ga.data_ptr()), ctypes.c_void_p(ge.data_ptr()), ctypes.c_void_p(s.cuda_stream))
    torch.cuda.synchronize()
    assert rc == 1 and bool((ga == 5.0).all()) and bool((ge == 5.0).all())
    # the library is still usable: the same renderer renders its frame
    f2 = g.render()
    torch.cuda.synchronize()
    assert float(f2.data[:, :3].max()) > 0
    with pytest.raises(D.DumbRayError):
        g.renderBlocksGPU(frame, 0, nb + 1, 1)


def test_cpp_mirror_reports_the_failure(dev):
    """The C++ mirror's renderBlocksGPU returns false with dr_last_error() set and the frame
    unchanged, which is where the reference's call site falls back to renderBlocksCPU."""
    exe = os.path.join(ROOT, "tests", "cpp", "host_api_demo")
    print("[case] C++ mirror failure path", flush=True)
    r = subprocess.run([exe, "--failure"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "renderBlocksGPU failed: render: material table smaller than the ids in use" in r.stdout
    assert "frame untouched" in r.stdout and "retry rendered" in r.stdout
