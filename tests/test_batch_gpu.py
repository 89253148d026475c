"""Batched uneven requests (dr_batch_*): one persistent launch replayed from a
hipGraph must give every request exactly the frame dr_render_blocks gives it
alone (same code path, same Philox keys), on every replay, with other GPU work
between the replays."""
import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S
from dumbray_amd.render import block_to_image_index

pytestmark = pytest.mark.gpu


def _renderers():
    scenes = [S.spheres_scene(8, seed=0), S.icosphere_scene(level=2, seed=1), S.spheres_scene(12, seed=4, n_lights=2)]
    dev = []
    for sc in scenes:
        t = D.Octree(16, 12)
        t.push(sc.tris)
        t.build()
        dev.append((sc, t.upload()))
    rng = np.random.default_rng(7)
    out = []
    for k in range(12):
        sc, ds = dev[k % 3]
        W, H = [(16, 16), (40, 24), (64, 64), (33, 17)][rng.integers(0, 4)]
        spp = [(1, 1), (2, 2), (4, 4)][rng.integers(0, 3)]
        depth = [2, 4, 8][rng.integers(0, 3)]
        r = D.DumbRenderer(ds, sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, spp, depth, rr_start=3,
                           rr_survive=0.8, seed=1000 + k)
        out.append((r, k + 1))
    return out


def _solo(r, req_id):
    import torch
    f = D.BlockBasedFrameBuffer(r.width, r.height)
    p = r.params(1, req_id)
    import ctypes
    from dumbray_amd._lib import check, lib
    check(lib().dr_render_blocks(r.scene.handle, ctypes.byref(p), ctypes.c_void_p(f.data.data_ptr()), 0,
                                 f.getBlockCount(), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "solo")
    torch.cuda.synchronize()
    return f.data.cpu().numpy()


@pytest.mark.parametrize("tuning", [None, {"tile_blocks": 3, "groups_per_cu": 2, "refill_min": 8}])
def test_batch_equals_individual_renders(dev, tuning):
    import torch
    rs = _renderers()
    ref = [_solo(r, rid) for r, rid in rs]
    frames = [D.BlockBasedFrameBuffer(r.width, r.height) for r, _ in rs]
    b = D.RenderBatch([(r, f, 1, rid) for (r, rid), f in zip(rs, frames)], tuning)
    t = b.tuning
    assert t["tile_blocks"] >= 1 and t["groups_per_cu"] >= 1 and t["refill_min"] >= 1
    print(f"[case] batch of {len(rs)} requests, tuning={t}", flush=True)
    for replay in range(3):
        for f in frames:
            f.data.fill_(-1.0)
        b.run()
        torch.cuda.synchronize()
        for k, f in enumerate(frames):
            got = f.data.cpu().numpy()
            inside = np.zeros(got.shape[0], bool)
            inside[block_to_image_index(f.width, f.height)] = True
            # pixels of the image: bit-identical to the individual render
            assert np.array_equal(got[inside].view(np.int32), ref[k][inside].view(np.int32)), (replay, k)
            # padding of partial blocks is never written (blockPixelLocation fails there)
            assert np.all(got[~inside] == -1.0), (replay, k)
        _solo(*rs[replay])          # eager work between replays


def _mixed_renderers(tmp_path):
    """Requests of both plugin sets in one batch: the reference's example scene (reference
    plugins) at several sizes and sample counts, a textured scene file, a spherical segment
    lens under each plugin set, and benchmark scenes."""
    import os
    from dumbray_amd import plugins as PL
    from test_scene_file import write_textured_scene
    here = os.path.dirname(os.path.abspath(__file__))
    ex = D.DumbRenderContext.fromFile(os.path.join(here, "golden", "example_scene_0", "0.dumb"))
    ex_tree = ex.octree()
    tx = D.DumbRenderContext.fromFile(write_textured_scene(tmp_path))
    tx_tree = tx.octree()
    out = [(ex.renderer(40, 24, tree=ex_tree, seed=11), 1), (ex.renderer(33, 17, tree=ex_tree, seed=12,
                                                                       samples_per_pixel=(2, 2)), 2),
           (ex.renderer(64, 48, tree=ex_tree, seed=13, max_bounces=3), 3),
           (tx.renderer(32, 32, tree=tx_tree, seed=14, samples_per_pixel=(2, 1)), 4)]
    lp = PL.spherical_segment_lens((70.0, 45.0), 12.0, 0.3, (1.0, 0.8, 0.6))
    out.append((ex.renderer(24, 40, tree=ex_tree, seed=15, lens=PL.SPHERICAL_SEGMENT_LENS, lens_params=lp,
                            cam_rows=ex.cam_rows * np.float32(1.25)), 5))
    sc = S.spheres_scene(8, seed=0)
    t = D.Octree(16, 12)
    t.push(sc.tris)
    t.build()
    ds = t.upload()
    rows = np.array(sc.camera.as_array()[3:12], np.float32)
    out.append((D.DumbRenderer(ds, sc.mat_id, sc.materials, sc.lights, sc.camera, 40, 30, (2, 2), 3, seed=16,
                               lens="spherical_segment", lens_params=lp, cam_rows=rows), 6))
    out.append((D.DumbRenderer(ds, sc.mat_id, sc.materials, sc.lights, sc.camera, 48, 16, (1, 1), 4, seed=17), 7))
    return out


@pytest.mark.parametrize("tuning", [None, {"tile_blocks": 2, "groups_per_cu": 1, "refill_min": 8}])
def test_mixed_plugin_batch_equals_individual_renders(dev, tmp_path, tuning):
    """The batch runs one persistent kernel per plugin set; every request still gets the
    frame dr_render_blocks gives it alone, on every replay."""
    import torch
    rs = _mixed_renderers(tmp_path)
    ref = [_solo(r, rid) for r, rid in rs]
    frames = [D.BlockBasedFrameBuffer(r.width, r.height) for r, _ in rs]
    b = D.RenderBatch([(r, f, 1, rid) for (r, rid), f in zip(rs, frames)], tuning)
    print(f"[case] mixed batch of {len(rs)} requests, tuning={b.tuning}", flush=True)
    for replay in range(3):
        for f in frames:
            f.data.fill_(-1.0)
        b.run()
        torch.cuda.synchronize()
        for k, f in enumerate(frames):
            got = f.data.cpu().numpy()
            inside = np.zeros(got.shape[0], bool)
            inside[block_to_image_index(f.width, f.height)] = True
            assert np.array_equal(got[inside].view(np.int32), ref[k][inside].view(np.int32)), (replay, k)
            assert np.all(got[~inside] == -1.0), (replay, k)
            assert ref[k][inside][:, :3].max() > 0, k
        _solo(*rs[replay])
