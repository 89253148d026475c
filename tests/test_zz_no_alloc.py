"""dr_backward_noalloc (the adjoint for callers that cannot allocate device memory) against
dr_backward through the C ABI: two calls in a row on one device, the second smaller, each
equal to dr_backward's gradients.  Both sum fp32 per-block partials into FP64 with atomics in
an order that depends on scheduling, so they agree to rounding, not bit for bit (dr_backward
against the FP64 oracle adjoint and finite differences is tested elsewhere)."""
import ctypes

import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import scenes as S
from dumbray_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def _grads(fn, g, dl):
    import torch
    ga = torch.zeros((g.nmat, 3), dtype=torch.float64, device="cuda")
    ge = torch.zeros((max(g.nlights, 1), 3), dtype=torch.float64, device="cuda")
    p = g.params(1, 0)
    check(fn(g.scene.handle, ctypes.byref(p), ctypes.c_void_p(dl.data_ptr()), ctypes.c_void_p(ga.data_ptr()),
             ctypes.c_void_p(ge.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "backward")
    torch.cuda.synchronize()
    return ga.cpu().numpy(), ge.cpu().numpy()[: g.nlights]


@pytest.mark.parametrize("order", ["large_then_small"])
def test_backward_noalloc_equals_backward(dev, order):
    import torch
    check(lib().dr_device_open(0), "dr_device_open")
    # the first call needs more vertex records than the array holds for every resident workgroup
    # (depth 8: 1024 blocks, 83 KB of records per workgroup), so its grid is the smaller one
    cases = [(S.spheres_scene(64, seed=5, n_lights=4, ground_material=False), 512, 256, (2, 2), 8),
             (S.icosphere_scene(level=2, seed=3), 40, 24, (1, 2), 2)]
    for k, (sc, W, H, spp, depth) in enumerate(cases):
        print(f"[case] noalloc call {k}: {W}x{H} spp {spp} depth {depth}, {len(sc.materials)} materials", flush=True)
        t = D.Octree(16, 12)
        t.push(sc.tris)
        t.build()
        g = D.DumbRenderer(t.upload(), sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, spp, depth, rr_start=3,
                           rr_survive=0.75)
        w = np.random.default_rng(10 + k).uniform(-1, 1, (H, W, 3)).astype(np.float32)
        dl = D.BlockBasedFrameBuffer(W, H).from_image(torch.from_numpy(w).cuda())
        na, ne = _grads(lib().dr_backward_noalloc, g, dl)
        ra, re = _grads(lib().dr_backward, g, dl)
        sa, se = np.abs(ra).max(), np.abs(re).max()
        assert sa > 0 and se > 0
        # fp32 per-block partial sums in scheduling order: rounding of the partials, relative to the
        # largest gradient (measured: 6.5e-8 of it at depth 8)
        assert np.all(np.abs(na - ra) <= 1e-5 * np.abs(ra) + 1e-6 * sa), np.abs(na - ra).max() / sa
        assert np.all(np.abs(ne - re) <= 1e-5 * np.abs(re) + 1e-6 * se), np.abs(ne - re).max() / se
