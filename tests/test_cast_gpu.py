"""GPU parity of the HIP closest-hit path (dr_cast through the C ABI) against the
oracle's Octree::cast on identical inputs: hit ids, t and hit points must be
bit-exact (integer/index work and fp32 arithmetic evaluated identically)."""
import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _rays(tris, n, seed, cam=None):
    inc = S.incoherent_rays(tris, n, seed=seed)
    lo, hi = S.scene_aabb(tris)
    rng = np.random.default_rng(seed)
    cen = tris.reshape(-1, 3, 3).mean(axis=1)
    pick = cen[rng.integers(0, len(cen), n // 2)]
    o = (lo + (hi - lo) * rng.uniform(-0.5, 1.5, (n // 2, 3))).astype(np.float32)
    aimed = np.concatenate([o, (pick - o).astype(np.float32)], axis=1)
    parts = [inc, aimed]
    if cam is not None:
        parts.append(S.primary_rays(cam, 64, 48))
    return np.ascontiguousarray(np.concatenate(parts, axis=0).astype(np.float32))


def _gpu_cast(tris, rays, clip, thresh=16, depth=12, exclude=None):
    import torch
    tree = D.Octree(thresh, depth)
    tree.push(tris)
    tree.build()
    sc = tree.upload()
    r = torch.from_numpy(rays).cuda()
    ex = None if exclude is None else torch.from_numpy(exclude).cuda()
    out = sc.cast(r, clip_backfaces=clip, exclude=ex, points=True)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _assert_same(g, o):
    bad = np.nonzero(g["id"] != o["id"])[0]
    assert len(bad) == 0, [(int(i), int(g["id"][i]), int(o["id"][i]), float(g["t"][i]), float(o["t"][i])) for i in bad[:5]]
    hit = o["id"] >= 0
    assert np.array_equal(g["t"][hit].view(np.int32), o["t"][hit].view(np.int32))
    assert np.array_equal(g["point"][hit].view(np.int32), o["point"][hit].view(np.int32))
    assert np.all(g["t"][~hit] == np.finfo(np.float32).max)


SCENES = {
    "soup": (lambda: S.random_soup(4000, seed=1, extent=10.0, size=1.5), None),
    "soup_dense": (lambda: S.random_soup(3000, seed=4, extent=3.0, size=2.0), None),
    "heightfield256": (lambda: S.heightfield(256, seed=1), None),
    "spheres": (lambda: S.spheres_scene(8, seed=0).tris, S.spheres_scene(8, seed=0).camera),
    "icosphere5": (lambda: S.icosphere((0, 0, 0), 1.0, 5, displace=0.05, seed=2), None),
}


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("clip", [False, True])
def test_gpu_cast_bit_exact(dev, oracle, name, clip):
    make, cam = SCENES[name]
    tris = make()
    rays = _rays(tris, 20000, seed=len(name), cam=cam)
    print(f"[case] {name} clip={clip} tris={len(tris)} rays={len(rays)}", flush=True)
    g = _gpu_cast(tris, rays, clip)
    o = oracle.Octree(tris, 16, 12).cast(rays, clip=clip, points=True)
    assert (o["id"] >= 0).mean() > 0.1
    _assert_same(g, o)


def test_gpu_cast_exclusion(dev, oracle):
    tris = S.spheres_scene(8, seed=0).tris
    rays = _rays(tris, 8000, seed=3)
    first = oracle.Octree(tris, 16, 12).cast(rays)
    ex = first["id"].copy()
    g = _gpu_cast(tris, rays, False, exclude=ex)
    o = oracle.Octree(tris, 16, 12).cast(rays, exclude=ex, points=True)
    _assert_same(g, o)


@pytest.mark.parametrize("thresh,depth", [(8, 24), (4, 8), (64, 2)])
def test_gpu_cast_knobs(dev, oracle, thresh, depth):
    tris = S.random_soup(2500, seed=7, extent=5.0, size=1.5)
    rays = _rays(tris, 10000, seed=thresh)
    g = _gpu_cast(tris, rays, False, thresh, depth)
    o = oracle.Octree(tris, thresh, depth).cast(rays, points=True)
    _assert_same(g, o)


def _reached_depth(tris, thresh, depth):
    t = D.Octree(thresh, depth)
    t.push(tris)
    t.build()
    _, ch, _, _ = t.export()
    d = np.zeros(len(ch), np.int64)
    for i in range(len(ch)):             # breadth-first order: parents precede their children
        if ch[i] >= 0:
            d[ch[i]:ch[i] + 8] = d[i] + 1
    return int(d.max())


def test_gpu_cast_stack_sized_by_reached_depth(dev, oracle):
    """The traversal stack holds one level per depth the build reached, not per level the
    depth limit allows: trees that stop far above their limit of 24 (a 32,768-tri
    heightfield split down to single triangles, depth 11; a sparse soup, depth 4) give the
    oracle's hits."""
    cases = ((S.heightfield(128, seed=1), 1, 11), (S.random_soup(2500, seed=7, extent=5.0, size=1.5), 8, 4))
    for tris, thresh, reached in cases:
        assert _reached_depth(tris, thresh, 24) == reached
        rays = _rays(tris, 8000, seed=21)
        print(f"[case] reached-depth stack: tris={len(tris)} thresh={thresh} reached={reached}", flush=True)
        ref = oracle.Octree(tris, thresh, 24).cast(rays, points=True)
        assert (ref["id"] >= 0).mean() > 0.1
        _assert_same(_gpu_cast(tris, rays, False, thresh, 24), ref)


def test_gpu_cast_edge_cases(dev, oracle):
    import torch
    # axis-aligned rays (zero direction components -> infinite inverse), rays through
    # vertices, a single-triangle scene (root leaf: no box test), an empty scene, n = 0
    tris = S.heightfield(16, seed=2)
    v = tris.reshape(-1, 3)[:200]
    rays = []
    for p in v:
        rays.append([p[0], p[1] + 50.0, p[2], 0.0, -1.0, 0.0])       # straight down onto a vertex
        rays.append([p[0] + 0.25, -50.0, p[2] + 0.5, 0.0, 1.0, 0.0])  # straight up
        rays.append([-5.0, p[1], p[2] + 0.3, 1.0, 0.0, 0.0])         # along +x
    rays = np.array(rays, np.float32)
    for clip in (False, True):
        _assert_same(_gpu_cast(tris, rays, clip), oracle.Octree(tris, 16, 12).cast(rays, clip=clip, points=True))
    one = S.random_soup(1, seed=5)
    r1 = _rays(one, 200, seed=1)
    _assert_same(_gpu_cast(one, r1, False), oracle.Octree(one, 16, 12).cast(r1, points=True))
    empty = np.zeros((0, 9), np.float32)
    g = _gpu_cast(empty, r1, False)
    assert np.all(g["id"] == -1)
    g0 = _gpu_cast(tris, np.zeros((0, 6), np.float32), False)
    assert g0["id"].shape == (0,)


def test_gpu_cast_full_heightfield(dev, oracle):
    """2,097,152-triangle heightfield with 4,194,304 incoherent rays + 1080p primaries:
    a 262,144-ray sample of each batch is compared bit-exactly with the oracle, and the
    whole batch is checked for size-independent properties (a hit point lies on its
    ray at t, misses carry FLT_MAX)."""
    import torch
    tris, coherent, incoherent, cam = S.raycast_workload()
    print(f"[case] full heightfield tris={len(tris)} rays={len(incoherent)}+{len(coherent)}", flush=True)
    tree = D.Octree(16, 12)
    tree.push(tris)
    tree.build()
    sc = tree.upload()
    otree = oracle.Octree(tris, 16, 12)
    rng = np.random.default_rng(0)
    for rays in (incoherent, coherent):
        r = torch.from_numpy(rays).cuda()
        out = sc.cast(r, clip_backfaces=False, points=True)
        torch.cuda.synchronize()
        g = {k: v.cpu().numpy() for k, v in out.items()}
        idx = np.sort(rng.choice(len(rays), 262144, replace=False))
        o = otree.cast(rays[idx], points=True)
        _assert_same({k: v[idx] for k, v in g.items()}, o)
        hit = g["id"] >= 0
        p = rays[hit, :3].astype(np.float64) + rays[hit, 3:].astype(np.float64) * g["t"][hit, None]
        assert np.max(np.abs(p - g["point"][hit])) < 1e-2
        assert np.all(g["t"][~hit] == np.finfo(np.float32).max)
        print(f"[case] batch n={len(rays)} hit={hit.mean():.3f}", flush=True)


def test_gpu_cast_concurrent_streams(dev, oracle):
    """Casts in flight on several streams at once (each launch takes its own pair of
    work counters from the scene's ring) give the oracle's results, on many launches in
    a row (the ring wraps) and on batches smaller than one wave's range of rays."""
    import torch
    tris = S.heightfield(128, seed=3)
    tree = D.Octree(16, 12)
    tree.push(tris)
    tree.build()
    sc = tree.upload()
    ot = oracle.Octree(tris, 16, 12)
    sizes = [5, 200, 7000, 60000]
    rays = [_rays(tris, n, seed=40 + k) for k, n in enumerate(sizes)]
    refs = [ot.cast(r, clip=False) for r in rays]
    dev_rays = [torch.from_numpy(r).cuda() for r in rays]
    streams = [torch.cuda.Stream() for _ in sizes]
    print(f"[case] concurrent casts on {len(streams)} streams, 300 rounds", flush=True)
    torch.cuda.synchronize()
    for rnd in range(300):                     # > the ring of 1024 counter pairs after all rounds
        outs = []
        for s, r in zip(streams, dev_rays):
            with torch.cuda.stream(s):
                outs.append(sc.cast(r, clip_backfaces=False, stream=s))
        torch.cuda.synchronize()
        if rnd % 50 == 0 or rnd == 299:
            for o, ref in zip(outs, refs):
                assert np.array_equal(o["id"].cpu().numpy(), ref["id"]), rnd
                assert np.array_equal(o["t"].cpu().numpy().view(np.int32), ref["t"].view(np.int32)), rnd
