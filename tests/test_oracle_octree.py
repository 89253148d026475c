"""Pins the oracle's octree build + cast (Octree.impl.cuh) against brute force:
the octree's closest hit must equal the closest hit over all triangles in push
order (strict '<', first wins) for every ray, on scenes of each kind, with and
without backface clipping and with the validateNotSameAsObject exclusion."""
import numpy as np
import pytest

from dumbray_amd import scenes as S

SCENES = {
    "soup": lambda: S.random_soup(3000, seed=1, extent=10.0, size=1.5),
    "soup_dense": lambda: S.random_soup(1500, seed=4, extent=3.0, size=2.0),
    "heightfield": lambda: S.heightfield(48, seed=1),
    "spheres": lambda: S.spheres_scene(8, seed=0).tris,
    "icosphere": lambda: S.icosphere((0, 0, 0), 1.0, 3, displace=0.05, seed=2),
}


def _rays(tris, n, seed):
    inc = S.incoherent_rays(tris, n, seed=seed)
    lo, hi = S.scene_aabb(tris)
    rng = np.random.default_rng(seed)
    # rays from outside aimed at triangle centroids (many exact-hit candidates)
    cen = tris.reshape(-1, 3, 3).mean(axis=1)
    pick = cen[rng.integers(0, len(cen), n // 2)]
    o = (lo + (hi - lo) * rng.uniform(-0.5, 1.5, (n // 2, 3))).astype(np.float32)
    d = (pick - o).astype(np.float32)
    aimed = np.concatenate([o, d], axis=1).astype(np.float32)
    return np.ascontiguousarray(np.concatenate([inc, aimed], axis=0))


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("clip", [False, True])
def test_octree_equals_brute_force(oracle, name, clip):
    tris = SCENES[name]()
    tree = oracle.Octree(tris, 16, 12)
    rays = _rays(tris, 4000, seed=hash(name) % 1000)
    a = tree.cast(rays, clip=clip)
    b = oracle.brute_cast(tris, rays, clip=clip)
    assert (a["id"] >= 0).mean() > 0.2
    bad = np.nonzero(a["id"] != b["id"])[0]
    assert len(bad) == 0, [(int(i), a["id"][i], b["id"][i], a["t"][i], b["t"][i]) for i in bad[:5]]
    assert np.array_equal(a["t"].view(np.int32), b["t"].view(np.int32))


@pytest.mark.parametrize("name", ["soup", "spheres"])
def test_exclusion_equals_brute_force(oracle, name):
    tris = SCENES[name]()
    tree = oracle.Octree(tris, 16, 12)
    rays = _rays(tris, 3000, seed=5)
    first = oracle.brute_cast(tris, rays)
    ex = first["id"].copy()
    a = tree.cast(rays, exclude=ex)
    b = oracle.brute_cast(tris, rays, exclude=ex)
    assert np.array_equal(a["id"], b["id"])
    assert np.all(a["id"][ex >= 0] != ex[ex >= 0])


@pytest.mark.parametrize("thresh,depth", [(8, 24), (16, 12), (4, 6), (64, 3)])
def test_knobs_and_structure(oracle, thresh, depth):
    tris = S.random_soup(1200, seed=9, extent=4.0, size=1.0)
    tree = oracle.Octree(tris, thresh, depth)
    bounds, children, off, lst = tree.export()
    n = len(children)
    assert (n - 1) % 8 == 0
    # depth of every node, children contiguous in groups of 8
    depthv = np.zeros(n, int)
    for i in range(n):
        if children[i] >= 0:
            assert children[i] + 7 < n
            depthv[children[i]:children[i] + 8] = depthv[i] + 1
            assert off[i + 1] == off[i]                       # interior nodes hold no list
    assert depthv.max() <= depth
    leaves = children < 0
    # leaves below max depth with >= thresh triangles exist only if the split was refused;
    # every non-degenerate triangle is in at least one leaf
    assert set(np.unique(lst)) == set(range(len(tris)))
    # cast agrees with brute force under these knobs too
    rays = _rays(tris, 2000, seed=thresh)
    assert np.array_equal(tree.cast(rays)["id"], oracle.brute_cast(tris, rays)["id"])
    assert leaves.sum() >= 1


def test_empty_and_single(oracle):
    rays = S.incoherent_rays(S.random_soup(4, seed=1), 100, seed=1)
    empty = oracle.Octree(np.zeros((0, 9), np.float32), 16, 12)
    assert np.all(empty.cast(rays)["id"] == -1)
    one = S.random_soup(1, seed=3)
    t1 = oracle.Octree(one, 16, 12)
    assert t1.node_count == 1
    r = np.concatenate([np.zeros((1, 3), np.float32), one.reshape(3, 3).mean(0, keepdims=True)], axis=1)
    assert t1.cast(r)["id"][0] == 0
