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
