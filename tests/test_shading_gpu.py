"""Shading tables per renderer (dr_shading) and the lifetime of what a batch refers to.

The reference keeps one material per Renderable element of the scene (Scene.cuh:68); the
C ABI lets several renderers of one device scene carry different material ids, vertex
normals and texture coordinates at the same time, and a batch keeps the scenes and
shading tables of its requests alive until it is destroyed."""
import ctypes

import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S
from dumbray_amd._lib import check, lib
from dumbray_amd.render import BatchTuning, Request

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def _tree(tris):
    t = D.Octree(16, 12)
    t.push(tris)
    t.build()
    return t


def _img(g, frame=None):
    import torch
    f = g.render(frame)
    torch.cuda.synchronize()
    return f.to_image()[..., :3].cpu().numpy()


def test_two_renderers_of_one_scene_with_different_material_ids(dev, oracle):
    import torch
    sc = S.spheres_scene(8, seed=0)
    t = _tree(sc.tris)
    ds = t.upload()
    W, H = 40, 30
    nmat = len(sc.materials)
    mid_b = ((sc.mat_id + 3) % nmat).astype(np.int32)         # another assignment of the same materials
    print("[case] two renderers, one scene, different material ids", flush=True)
    a = D.DumbRenderer(ds, sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, (2, 2), 3)
    b = D.DumbRenderer(ds, mid_b, sc.materials, sc.lights, sc.camera, W, H, (2, 2), 3)
    assert a.shading is not b.shading
    img_a, img_b = _img(a), _img(b)
    ot = oracle.Octree(sc.tris, 16, 12)
    for img, mid in ((img_a, sc.mat_id), (img_b, mid_b)):
        ref = oracle.Renderer(ot, mid, sc.materials, sc.lights, sc.camera.as_array(), W, H, (2, 2), 3).render()
        err = np.abs(img - ref)
        assert ref.max() > 0 and not (err > RTOL * np.abs(ref) + ATOL).any()
    assert not np.array_equal(img_a, img_b)
    # renderers with equal tables share one handle
    c = D.DumbRenderer(ds, sc.mat_id.copy(), sc.materials, sc.lights, sc.camera, W, H, (1, 1), 2)
    assert c.shading is a.shading
    # a batch of a and b, then a renderer with yet other ids on the same scene: replays unchanged
    fa, fb = D.BlockBasedFrameBuffer(W, H), D.BlockBasedFrameBuffer(W, H)
    batch = D.RenderBatch([(a, fa, 1, 0), (b, fb, 1, 0)], {"tile_blocks": 1, "groups_per_cu": 1, "refill_min": 8})
    d = D.DumbRenderer(ds, np.zeros_like(sc.mat_id), sc.materials, sc.lights, sc.camera, W, H, (2, 2), 3)
    _img(d)
    for replay in range(2):
        fa.data.fill_(-1.0)
        fb.data.fill_(-1.0)
        batch.run()
        torch.cuda.synchronize()
        assert np.array_equal(fa.to_image()[..., :3].cpu().numpy(), img_a), replay
        assert np.array_equal(fb.to_image()[..., :3].cpu().numpy(), img_b), replay


def test_batch_keeps_destroyed_scene_alive(dev):
    """dr_scene_destroy and dr_shading_destroy while a batch refers to them: the batch still
    renders the request's frame (the memory goes with dr_batch_destroy)."""
    import torch
    sc = S.spheres_scene(8, seed=2)
    t = _tree(sc.tris)
    W, H = 33, 20
    h = ctypes.c_void_p()
    check(lib().dr_scene_create(t._h, ctypes.byref(h)), "dr_scene_create")
    sh = ctypes.c_void_p()
    mid = np.ascontiguousarray(sc.mat_id, np.int32)
    check(lib().dr_shading_create(h, mid.ctypes.data_as(ctypes.c_void_p), None, None, ctypes.byref(sh)),
          "dr_shading_create")
    # a renderer on a scene of its own gives the expected frame
    g = D.DumbRenderer(t.upload(), sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, (2, 2), 3)
    want = _img(g)
    p = g.params(1, 0)
    p.shading = sh.value
    frame = D.BlockBasedFrameBuffer(W, H)
    reqs = (Request * 1)()
    reqs[0].scene = h.value
    reqs[0].params = p
    reqs[0].frame = frame.data.data_ptr()
    tun = BatchTuning(1, 1, 8)
    bh = ctypes.c_void_p()
    check(lib().dr_batch_create(reqs, 1, ctypes.byref(tun), ctypes.byref(bh)), "dr_batch_create")
    lib().dr_shading_destroy(sh)
    lib().dr_scene_destroy(h)
    print("[case] batch replay after dr_scene_destroy / dr_shading_destroy", flush=True)
    try:
        for _ in range(2):
            frame.data.fill_(-1.0)
            check(lib().dr_batch_run(bh, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "dr_batch_run")
            torch.cuda.synchronize()
            assert np.array_equal(frame.to_image()[..., :3].cpu().numpy(), want)
    finally:
        lib().dr_batch_destroy(bh)


def test_shading_of_another_scene_is_refused(dev):
    sc = S.spheres_scene(8, seed=0)
    s1, s2 = _tree(sc.tris).upload(), _tree(sc.tris).upload()
    g = D.DumbRenderer(s1, sc.mat_id, sc.materials, sc.lights, sc.camera, 16, 16)
    p = g.params(1, 0)
    other = s2.shading(sc.mat_id)            # held for the call
    p.shading = other.handle.value
    f = D.BlockBasedFrameBuffer(16, 16)
    rc = lib().dr_render_blocks(s1.handle, ctypes.byref(p), ctypes.c_void_p(f.data.data_ptr()), 0, f.getBlockCount(),
                                None)
    assert rc == 1 and b"another scene" in lib().dr_last_error()
