"""GPU parity at the benchmark workloads themselves (full sizes, not scaled-down scenes):

* the 1920x1080, 16 spp, depth 8 path trace of three displaced level-7 icospheres
  (983,042 tris; Lambert, mirror, Schlick metal; two area lights; Russian roulette from
  depth 3) against the oracle estimator on a band of full rows (DumbRenderer.cu:87-113
  renders whole frames the same way, block by block);
* the 64-request batch over 8 scenes of 61k-983k tris: every request's batch frame equals
  its own dr_render_blocks frame bit for bit on two graph replays (BlockRenderer.cu:74-97
  drains the same blocks through host round trips), and one request per scene agrees with
  the oracle on a band of rows;
* the 512x512, 16 spp, depth 4 adjoint of 64 spheres with 64 materials and 4 emitters:
  the L2 loss against a target rendered from hidden parameters, dL/dalbedo of the 8
  materials with the largest gradients and dL/demission of all 4 emitters against central
  differences of GPU renders (eps 1e-3, fixed seeds, 1e-3 relative), and the whole gradient
  against the FP64 oracle adjoint with dL/dI nonzero on a 64x64 window.

Per-pixel tolerance against the oracle: |gpu - oracle| <= 1e-4 |oracle| + 1e-6 (the GPU
accumulates radiance forwards, the oracle bottom-up like the reference, which reassociates
fp32 sums; the paths themselves are identical)."""
import threading

import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S
from dumbray_amd.render import block_to_image_index

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def _close(img, ref, what):
    err = np.abs(img.astype(np.float64) - ref)
    bad = err > RTOL * np.abs(ref) + ATOL
    assert not bad.any(), (what, int(bad.sum()), float(err.max()), np.argwhere(bad)[:5].tolist())


# octree knobs of the 1080p and batch scenes: the reference's defaults (Octree.cuh:11,13)
REF_KNOBS = (8, 24)


def _device_scene(tris, split=16, depth=12):
    t = D.Octree(split, depth)
    t.push(tris)
    t.build()
    return t.upload()


def _oracle_trees(oracle, tri_sets, split=16, depth=12):
    """Oracle octrees of several scenes, built in parallel threads (the builds release the GIL)."""
    out = [None] * len(tri_sets)

    def work(k):
        out[k] = oracle.Octree(tri_sets[k], split, depth)
    th = [threading.Thread(target=work, args=(k,)) for k in range(len(tri_sets))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return out


def test_path_trace_1080p_band_matches_oracle(dev, oracle):
    import torch
    sc = S.icosphere_scene(level=7, seed=3)
    W, H, spp, depth = 1920, 1080, (4, 4), 8
    print(f"[case] 1080p 16 spp depth 8, {len(sc.tris)} tris", flush=True)
    g = D.DumbRenderer(_device_scene(sc.tris, *REF_KNOBS), sc.mat_id, sc.materials, sc.lights, sc.camera, W, H,
                       spp, depth, rr_start=3, rr_survive=0.75)
    frame = g.render()
    torch.cuda.synchronize()
    img = frame.to_image()[..., :3].cpu().numpy()
    o = oracle.Renderer(oracle.Octree(sc.tris, *REF_KNOBS), sc.mat_id, sc.materials, sc.lights,
                        sc.camera.as_array(), W, H, spp, depth, rr_start=3, rr_survive=0.75)
    assert np.isfinite(img).all() and img.max() > 0
    lit = np.where(img.max(axis=(1, 2)) > 0)[0]          # rows with geometry in view
    for r0 in (int(lit[len(lit) // 2]) - 8, int(lit[(3 * len(lit)) // 4]) - 8):
        ref = o.render(rows=(r0, r0 + 16))
        assert ref.max() > 0
        _close(img[r0:r0 + 16], ref, f"rows {r0}..{r0 + 15}")


def _batch_renderers():
    scenes, reqs = S.batch_workload()
    dev_scenes = [_device_scene(sc.tris, *REF_KNOBS) for sc in scenes]
    rs = []
    for q in reqs:
        sc = scenes[q["scene"]]
        r = D.DumbRenderer(dev_scenes[q["scene"]], sc.mat_id, sc.materials, sc.lights, sc.camera, q["width"],
                           q["height"], q["fsaa"], q["depth"], rr_start=3, rr_survive=0.75, seed=q["seed"])
        rs.append((r, q))
    return scenes, rs


def test_batch_64_requests_equal_solo_renders_and_oracle(dev, oracle):
    import torch
    scenes, rs = _batch_renderers()
    print(f"[case] 64-request batch over {len(scenes)} scenes ({sum(len(s.tris) for s in scenes)} tris)", flush=True)
    import ctypes
    from dumbray_amd._lib import check, lib
    solo = []
    for r, q in rs:
        # the request alone, with the Philox counter word the batch gives it
        f = D.BlockBasedFrameBuffer(r.width, r.height)
        p = r.params(1, q["request"])
        check(lib().dr_render_blocks(r.scene.handle, ctypes.byref(p), ctypes.c_void_p(f.data.data_ptr()), 0,
                                     f.getBlockCount(), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
              "dr_render_blocks")
        solo.append(f)
    torch.cuda.synchronize()
    ref = [f.data.cpu().numpy() for f in solo]
    frames = [D.BlockBasedFrameBuffer(r.width, r.height) for r, _ in rs]
    batch = D.RenderBatch([(r, f, 1, q["request"]) for (r, q), f in zip(rs, frames)])
    print(f"[case] batch tuning {batch.tuning}", flush=True)
    for replay in range(2):
        for f in frames:
            f.data.fill_(-1.0)
        batch.run()
        torch.cuda.synchronize()
        for k, f in enumerate(frames):
            got = f.data.cpu().numpy()
            inside = np.zeros(got.shape[0], bool)
            inside[block_to_image_index(f.width, f.height)] = True
            assert np.array_equal(got[inside].view(np.int32), ref[k][inside].view(np.int32)), (replay, k)
            assert np.all(got[~inside] == -1.0), (replay, k)
        print(f"[case] replay {replay}: 64 frames bit-identical to solo renders", flush=True)
    # one request per scene against the oracle estimator on a band of 16 rows
    trees = _oracle_trees(oracle, [sc.tris for sc in scenes], *REF_KNOBS)
    checked = set()
    for (r, q), f in zip(rs, frames):
        if q["scene"] in checked:
            continue
        checked.add(q["scene"])
        sc = scenes[q["scene"]]
        o = oracle.Renderer(trees[q["scene"]], sc.mat_id, sc.materials, sc.lights, sc.camera.as_array(), q["width"],
                            q["height"], q["fsaa"], q["depth"], rr_start=3, rr_survive=0.75, seed=q["seed"])
        r0 = q["height"] // 2 - 8
        band = o.render(rows=(r0, r0 + 16), request=q["request"])
        img = f.to_image()[..., :3].cpu().numpy()[r0:r0 + 16]
        assert band.max() > 0
        _close(img, band, f"request {q['request']} rows {r0}..{r0 + 15}")
    assert len(checked) == len(scenes)


def test_adjoint_512_matches_fd_and_fp64_window(dev, oracle):
    import torch
    sc = S.spheres_scene(64, seed=5, n_lights=4, ground_material=False)
    W, H = 512, 512
    g = D.DumbRenderer(_device_scene(sc.tris), sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, (4, 4), 4)
    print(f"[case] adjoint 512x512 16 spp depth 4, {len(sc.tris)} tris, {g.nmat} materials, {g.nlights} lights",
          flush=True)
    assert g.nmat == 64 and g.nlights == 4
    target = g.render().data[:, :3].clone()                 # hidden parameters
    rng = np.random.default_rng(1234)
    g.albedo.copy_(torch.from_numpy(rng.uniform(0.2, 0.8, (64, 3)).astype(np.float32)))
    g.emission.copy_(torch.from_numpy(rng.uniform(5.0, 15.0, (4, 3)).astype(np.float32)))

    def loss():
        img = g.render().data[:, :3]
        return float(((img.double() - target.double()) ** 2).sum())

    frame = g.render()
    dl = torch.zeros_like(frame.data)
    dl[:, :3] = 2.0 * (frame.data[:, :3] - target)
    ga, ge = g.backward(dl)
    torch.cuda.synchronize()
    ga, ge = ga.cpu().numpy(), ge.cpu().numpy()
    eps = 1e-3

    def fd(table, m, c, eps):
        base = float(table[m, c])
        up, dn = np.float32(base + eps), np.float32(base - eps)
        table[m, c] = float(up)
        lp = loss()
        table[m, c] = float(dn)
        lm = loss()
        table[m, c] = base
        return (lp - lm) / (float(up) - float(dn))
    top = np.argsort(-np.abs(ga).max(axis=1))[:8]
    for m in top:
        for c in range(3):
            v = fd(g.albedo, m, c, eps)
            assert abs(v - ga[m, c]) <= 1e-3 * abs(ga[m, c]), ("albedo", m, c, v, ga[m, c])
    # The image is linear in emission, so a central difference has no truncation error and
    # only the fp32 rounding of the two renders limits it; the step is 1e-3 relative to the
    # emission (~1e-2) rather than absolute, which keeps that rounding 10x below the tolerance.
    for l in range(4):
        for c in range(3):
            v = fd(g.emission, l, c, eps * abs(float(g.emission[l, c])))
            assert abs(v - ge[l, c]) <= 1e-3 * abs(ge[l, c]), ("emission", l, c, v, ge[l, c])
    print(f"[case] FD: 8 albedo rows (|g| {np.abs(ga[top]).min():.3e}..{np.abs(ga[top]).max():.3e}) and 4 emitters "
          f"within 1e-3", flush=True)
    # FP64 oracle adjoint of the same estimator, dL/dI nonzero on a 64x64 window
    win = frame.to_image()[..., :3].cpu().numpy().astype(np.float64)
    tgt = D.BlockBasedFrameBuffer(W, H)
    tgt.data[:, :3] = target
    timg = tgt.to_image()[..., :3].cpu().numpy().astype(np.float64)
    w = np.zeros((H, W, 3), np.float32)
    w[224:288, 224:288] = (2.0 * (win - timg))[224:288, 224:288]
    dlw = frame.from_image(torch.from_numpy(w).cuda())
    gaw, gew = g.backward(dlw)
    torch.cuda.synchronize()
    gaw, gew = gaw.cpu().numpy(), gew.cpu().numpy()
    o = oracle.Renderer(oracle.Octree(sc.tris, 16, 12), sc.mat_id, sc.materials, sc.lights, sc.camera.as_array(),
                        W, H, (4, 4), 4)
    for m in range(64):
        o.set_albedo(m, g.albedo[m].cpu().numpy())
    for l in range(4):
        o.set_emission(l, g.emission[l].cpu().numpy())
    oa, oe = o.backward(w)
    scale = np.abs(oa).max()
    assert scale > 0 and np.abs(oe).max() > 0
    assert np.all(np.abs(gaw - oa) <= 1e-4 * np.abs(oa) + 1e-6 * scale), np.abs(gaw - oa).max()
    assert np.all(np.abs(gew - oe) <= 1e-4 * np.abs(oe) + 1e-9), (gew, oe)
