"""GPU renders of scene files against the oracle: the reference's example scene (cargo
boxes: DumbBased / SimpleStochastic materials, soft directional and spherical lights,
OBJ vertex normals) and a synthetic scene with groups, includes and non-default
renderer settings, loaded by the library's C++ loader for the GPU and by the oracle's
own restatement for the oracle render.

Per pixel |gpu - oracle| <= 1e-4 * |oracle| + 1e-6 (the GPU accumulates path
throughput forward, the oracle layer colours bottom-up; everything else is the same
fp32 arithmetic)."""
import os
import textwrap

import numpy as np
import pytest

import dumbray_amd as D
import oracle
from oracle import scene_file as SF

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLE = os.path.join(HERE, "golden", "example_scene_0", "0.dumb")
RTOL, ATOL = 1e-4, 1e-6


def _close(a, b):
    err = np.abs(a - b)
    bad = err > RTOL * np.abs(b) + ATOL
    assert not bad.any(), (int(bad.sum()), float(err.max()), np.argwhere(bad)[:5].tolist())


def _pair(path, W, H, iterations=1, **over):
    import torch

    ctx = D.DumbRenderContext.fromFile(path)
    sf = SF.load(path)
    g = ctx.renderer(W, H, **over)
    boxing = {"height": 0, "width": 1, "min": 2, "max": 3}[over.get("boxing", ctx.boxing)]
    o = oracle.Renderer(oracle.Octree(sf.tris, 8, 24), sf.mat_id, sf.materials, sf.lights, sf.cam, W, H,
                        over.get("samples_per_pixel", sf.fsaa), over.get("max_bounces", sf.max_bounces),
                        ignore_backfaces=sf.ignore_backfaces, normals=sf.normals, boxing=boxing, lens=sf.lens,
                        lens_params=sf.lens_params, cam_rows=sf.cam_rows,
                        textures=sf.textures if sf.textures else None, uv=sf.uv)
    frame = None
    for it in range(1, iterations + 1):
        print(f"[case] {os.path.basename(path)} {W}x{H} iteration {it} {over}", flush=True)
        frame = g.render(frame, it)
        torch.cuda.synchronize()
        img = frame.to_image()[..., :3].cpu().numpy()
        ref = o.render_ref(iteration=it, nthreads=8)
        assert ref.max() > 0
        if it == 1:
            _close(img, ref)
    return g, o


def test_example_scene_matches_oracle(dev):
    g, _ = _pair(EXAMPLE, 96, 54)
    assert g.plugins == D.plugins.PLUGINS_REFERENCE and g.max_bounces == 2


@pytest.mark.parametrize("boxing", ["width", "min", "max"])
def test_example_scene_boxing_modes_match_oracle(dev, boxing):
    _pair(EXAMPLE, 64, 40, boxing=boxing, samples_per_pixel=(2, 1))


def test_example_scene_more_bounces_and_samples(dev):
    _pair(EXAMPLE, 48, 32, samples_per_pixel=(2, 2), max_bounces=4)


SCENE = """
{ "include": ["inc.dumb"],
  "materials": [ {"name": "glossy", "type": "dumb_based", "preset": "glossy", "diffuse": [0.6, 0.3, 0.2]},
                 {"name": "stoch", "type": "simple_stochastic", "color": [0.3, 0.6, 0.9], "smoothness": 0.5} ],
  "lights": [ {"type": "simple_soft_directional", "direction": [-1, -2, 0.5], "color": [0.9, 0.8, 0.7],
               "softness": 0.05},
              {"type": "simple_spherical", "color": [60, 60, 50], "position": [-3, 9, -2], "radius": 1.0} ],
  "groups": [ {"name": "pair", "elements": [
                 {"mesh": {"obj": "mesh.obj", "object": "floor"}, "material": "glossy"},
                 {"mesh": {"obj": "mesh.obj", "object": "block"}, "transform": {"rotation": [0, 30, 0]}} ],
               "fallback_material": "stoch"} ],
  "objects": [ "pair",
               {"group": "pair", "override_material": "inc_mat",
                "transform": {"position": [2.5, 1.0, 1.5], "scale": [0.5, 0.5, 0.5], "rotation": [0, -20, 10]}} ],
  "camera": {"lense": {"type": "simple_stochastic", "angle": 55},
             "transform": {"position": [0, 1, 0], "rotation": [25, 10, 0], "relative_position": [0, 0, -9]}},
  "renderer": {"pixel": {"boxing": "max", "bounces": 3, "samples_per_pixel": [2, 2], "ignore_backfaces": false}} }
"""
INC = '{"materials": [{"name": "inc_mat", "type": "fresnel_color", "color": [0.8, 0.7, 0.2]}]}'
MESH = """v -5 0 -5
v 5 0 -5
v 5 0 5
v -5 0 5
v -1 0 -1
v 1 0 -1
v 1 0 1
v -1 0 1
v -1 2 -1
v 1 2 -1
v 1 2 1
v -1 2 1
vn 0 1 0
g floor
f 1//1 4//1 3//1 2//1
g block
f 5 6 7 8
f 9 12 11 10
f 5 9 10 6
f 6 10 11 7
f 7 11 12 8
f 8 12 9 5
"""


def test_grouped_scene_matches_oracle(dev, tmp_path):
    for name, text in (("scene.dumb", SCENE), ("inc.dumb", INC), ("mesh.obj", MESH)):
        (tmp_path / name).write_text(textwrap.dedent(text))
    g, _ = _pair(str(tmp_path / "scene.dumb"), 40, 32, iterations=2)
    assert g.boxing == 3 and g.fsaa_x == 2 and not g.ignore_backfaces


def test_example_scene_spherical_segment_lens_matches_oracle(dev, tmp_path):
    """The example scene seen through the spherical segment lens (angles, aperture, filter)."""
    import re

    with open(EXAMPLE) as f:
        text = f.read()
    text = text.replace('"files/cargo.obj"', '"%s"' % os.path.join(os.path.dirname(EXAMPLE), "files", "cargo.obj"))
    text = re.sub(r'"lense"\s*:\s*\{[^}]*\}',
                  '"lense": {"type": "spherical_segment", "angle": 40, "distance": 140, "sensor": 2.5, '
                  '"filter": [1.0, 0.9, 0.7]}', text)
    path = tmp_path / "lens.dumb"
    path.write_text(text)
    g, _ = _pair(str(path), 64, 40)
    assert g.lens == D.plugins.SPHERICAL_SEGMENT_LENS


def test_textured_scene_matches_oracle(dev, tmp_path):
    """PNG textures through the scene file: a nearest-filtered checker diffuse map with a normal
    map, an alpha-cut leaf with a fresnel-colour texture, a textured FresnelShader."""
    from test_scene_file import write_textured_scene

    path = write_textured_scene(tmp_path)
    g, _ = _pair(path, 48, 32, iterations=2)
    assert g.ntextures == 3 and g.nmaps == 3
