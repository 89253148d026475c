"""Writes spheres8_256_fp64_4096spp.npy: the config-0 scene (ground quad + 8 UV spheres,
scenes.spheres_scene(8, seed=0), one 1x1 area light at y = 5, emission 10) rendered at
256x256, max depth 4, by the oracle's FP64 path estimator at 4096 samples per pixel
(16 iterations of 16x16 subpixels, Philox key 99, independent of the GPU renders' keys).
Row-major (256, 256, 3) float32.  About a minute on 8 cores.

    python tests/golden/make_fp64_reference.py            # rewrites the fixture
    python tests/golden/make_fp64_reference.py rows 96 112  # prints a band (for the test)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spheres8_256_fp64_4096spp.npy")
W = H = 256
DEPTH, SEED, SUB, ITERS = 4, 99, (16, 16), 16


def render(rows=None, nthreads=0):
    import oracle
    from dumbray_amd import scenes as S
    oracle.build()
    sc = S.spheres_scene(8, seed=0)
    o = oracle.Renderer(oracle.Octree(sc.tris, 16, 12), sc.mat_id, sc.materials, sc.lights, sc.camera.as_array(),
                        W, H, SUB, DEPTH, seed=SEED)
    acc = None
    for k in range(1, ITERS + 1):
        img = o.render(iteration=k, fp64=True, rows=rows, nthreads=nthreads)
        acc = img if acc is None else acc + img
    return acc / ITERS


if __name__ == "__main__":
    if len(sys.argv) == 4 and sys.argv[1] == "rows":
        print(render(rows=(int(sys.argv[2]), int(sys.argv[3]))).tolist())
    else:
        np.save(OUT, render().astype(np.float32))
        print("wrote", OUT)
