"""The product's parallel host builder (dr_octree_build, C++) must produce the
oracle's tree exactly: the same boxes bit for bit, the same leaf lists in the
same stored order, the same child order.  Node indices may differ (the
reference's thread-pool build permutes them too), so trees are compared by a
canonical depth-first walk.  Host-only: runs without a GPU."""
import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S


def canon(bounds, children, off, lst):
    out, stack = [], [0]
    while stack:
        j = stack.pop()
        c = children[j]
        out.append((bounds[j].tobytes(), c >= 0, tuple(lst[off[j]:off[j + 1]].tolist())))
        if c >= 0:
            stack.extend(range(c + 7, c - 1, -1))
    return out


CASES = [
    ("soup", lambda: S.random_soup(3000, seed=1, extent=10.0, size=1.5), 16, 12),
    ("soup_t8_d24", lambda: S.random_soup(2000, seed=2, extent=5.0, size=2.0), 8, 24),
    ("heightfield128", lambda: S.heightfield(128, seed=1), 16, 12),
    # several nodes above the builder's team-split size (16384 triangles) at depths 0-2
    ("soup80k_t8_d24", lambda: S.random_soup(80000, seed=3, extent=20.0, size=1.0), 8, 24),
    ("spheres", lambda: S.spheres_scene(8, seed=0).tris, 16, 12),
    ("icosphere4", lambda: S.icosphere((0.5, -0.2, 0.1), 1.3, 4, displace=0.05, seed=1), 16, 12),
    ("single", lambda: S.random_soup(1, seed=1), 16, 12),
    ("empty", lambda: np.zeros((0, 9), np.float32), 16, 12),
]


@pytest.mark.parametrize("name,make,thresh,depth", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("nthreads", [1, 5])
def test_product_tree_equals_oracle(oracle, name, make, thresh, depth, nthreads):
    tris = make()
    o = oracle.Octree(tris, thresh, depth)
    p = D.Octree(thresh, depth)
    p.push(tris)
    p.build(nthreads=nthreads)
    assert p.getNodeCount() == o.node_count
    assert canon(*p.export()) == canon(*o.export())


def test_push_order_is_hit_id():
    a = S.random_soup(10, seed=1)
    b = S.random_soup(7, seed=2)
    p = D.Octree(16, 12)
    p.push(a)
    p.push(b)
    p.build()
    assert np.array_equal(p.getData(), np.concatenate([a, b]))


def test_bad_knobs_raise():
    p = D.Octree(16, 25)
    p.push(S.random_soup(5, seed=1))
    with pytest.raises(D.DumbRayError):
        p.build()


@pytest.mark.parametrize("name,make,thresh,depth", CASES[:5], ids=[c[0] for c in CASES[:5]])
def test_interior_children_own_consecutive_blocks(name, make, thresh, depth):
    """The device group layout finds an interior child's group as the group of the
    node's first interior child plus its rank: breadth-first export guarantees it."""
    p = D.Octree(thresh, depth)
    p.push(make())
    p.build(nthreads=3)
    _, children, _, _ = p.export()
    for node in np.flatnonzero(children >= 0):
        c = children[node]
        sub = children[c:c + 8]
        blocks = sub[sub >= 0]
        if len(blocks):
            assert np.array_equal(blocks, blocks[0] + 8 * np.arange(len(blocks))), node
