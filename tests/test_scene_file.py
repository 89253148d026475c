"""Scene files (.dumb + OBJ): the oracle restatement pinned by closed forms and by the
reference's own parser test inputs, then the library's C++ loader (dr_scene_file_*)
checked against the oracle value for value (bit-exact fp32) on the reference's example
scene and on synthetic scenes that exercise every section."""
import math
import os
import textwrap

import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import plugins as PL
import oracle
from oracle import scene_file as SF

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLE = os.path.join(HERE, "golden", "example_scene_0", "0.dumb")
REF_EXAMPLE = "/root/reference/example_scenes/0"
f32 = np.float32


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


# ----------------------------------------------------------------------------- oracle pins
DSON_GAUNTLET = [   # Dson.test.cu:447-470, outcomes read off Dson.cu:292-429
    ("", None), ("{}", {}), ("[]", []), ("{[]}", None), ("[{}]", [{}]), ("[ {}, {}]", [{}, {}]),
    ("[ {  } , {  } ]", [{}, {}]), ("[{},{}]", [{}, {}]), ('["aaa"]', ["aaa"]), ('["aaa",]', None),
    ('["aaa","bbb"]', ["aaa", "bbb"]), ('{"aaa"}', None), ('{"aaa":}', None), ('{"aaa":"bbb"}', {"aaa": "bbb"}),
    ('{"aaa":"bbb", "ccc"}', None), ('{"aaa":"bbb", "ccc" :}', None),
    ('{"aaa":"bbb", "ccc" : "ddd"}', {"aaa": "bbb", "ccc": "ddd"}),
    ('{"aaa":"bbb", "ccc" : "ddd",  "aaa":[]}', None),
    ("1", 1.0), ("-1", -1.0), ("1.21", 1.21), ("1junk", 1.0), ("some_garbage", None),
    ('{"a":true, "b": false, "c" : null, "d": 0, "e": [], "f": {}, "g": "text", "h": {"1":1}, "i":[0], '
     '"j" : {"1":"1", "2":[{},{}]}}',
     {"a": True, "b": False, "c": None, "d": 0.0, "e": [], "f": {}, "g": "text", "h": {"1": 1.0}, "i": [0.0],
      "j": {"1": "1", "2": [{}, {}]}}),
]


@pytest.mark.parametrize("text,expected", DSON_GAUNTLET)
def test_oracle_dson_gauntlet(text, expected):
    if expected is None:
        with pytest.raises(SF.SceneError):
            SF.parse_dson(text)
    else:
        assert SF.parse_dson(text) == expected


def test_oracle_dson_strings_numbers():
    v = SF.parse_dson(r'{"s": "a\"b\\c\/d\te\qx", "n": [1e3, -2.5E-1, +7, .5, 3.]}')
    assert v["s"] == 'a"b\\c/d\te\\qx'
    assert v["n"] == [1000.0, -0.25, 7.0, 0.5, 3.0]
    for bad in ("1e", "1e+", "-", ".", "inf", "nan"):
        with pytest.raises(SF.SceneError):
            SF.parse_dson(bad)


def test_oracle_transform_without_rotation_is_exact_scale_and_shift():
    """Zero Euler angles leave the unit rows exact, so a point maps to fl(fl(v*s) + p)."""
    rng = np.random.default_rng(1)
    t = SF.transform_from(SF.parse_dson('{"position": [3.25, -7, 0.1], "scale": [0.5, 3, -1.7]}'))
    p, s = np.array(t.position, f32), np.array(t.scale, f32)
    for v in rng.normal(0, 40, (50, 3)).astype(f32):
        got = np.array(t.point(tuple(v)), f32)
        assert same(got, (v * s) + p)
        assert same(np.array(t.normal(tuple(v)), f32), ((v * s) + p) - p)


def test_oracle_relative_position_without_rotation_adds_offset():
    t = SF.transform_from(SF.parse_dson('{"position": [1, 2, 3], "scale": [2, 4, 8], "relative_position": [5, -6, 7]}'))
    assert same(np.array(t.position, f32), np.array([6, -4, 10], f32))


def _rows64(euler_deg):
    """float64 statement of the Euler convention: rows start as the unit axes; turn x, z about
    y by angle y, then y, z about (the turned) x by angle x, then x, y about z by angle z, each
    v <- v cos a + (axis x v) sin a."""
    x, y, z = np.eye(3)

    def turn(axis, v, a):
        return v * math.cos(a) + np.cross(axis, v) * math.sin(a)

    ex, ey, ez = (math.radians(a) for a in euler_deg)
    x, z = turn(y, x, ey), turn(y, z, ey)
    y, z = turn(x, y, ex), turn(x, z, ex)
    x, y = turn(z, x, ez), turn(z, y, ez)
    return np.stack([x, y, z])


@pytest.mark.parametrize("euler", [(0, 90, 0), (90, 0, 0), (0, 0, 90), (48, 16, 0), (4, 32, 8), (-130, 75, 210)])
def test_oracle_rotation_matches_closed_form(euler):
    t = SF.Transform()
    t.euler = tuple(f32(a) for a in euler)
    t.rebake()
    rows = np.array([t.tx, t.ty, t.tz], np.float64)
    np.testing.assert_allclose(rows, _rows64(euler), atol=2e-6)
    np.testing.assert_allclose(rows @ rows.T, np.eye(3), atol=4e-6)
    assert abs(np.linalg.det(rows) - 1.0) < 4e-6


def test_oracle_rotation_y90_direction():
    """Rotation (0, 90, 0) sends +x to -z and +z to +x (rows (c, 0, -s), (0, 1, 0), (s, 0, c))."""
    t = SF.Transform()
    t.euler = (f32(0), f32(90), f32(0))
    t.rebake()
    np.testing.assert_allclose(t.point(SF.V(1, 0, 0)), (0, 0, 1), atol=1e-6)
    np.testing.assert_allclose(t.point(SF.V(0, 0, 1)), (-1, 0, 0), atol=1e-6)


def test_oracle_lens_focal():
    assert abs(float(SF.lens_focal(90.0)) - 0.5) < 1e-7
    assert abs(float(SF.lens_focal(60.0)) - 0.5 / math.tan(math.radians(30))) < 1e-6
    assert SF.lens_focal(0.5) == SF.lens_focal(1.0) and SF.lens_focal(500) == SF.lens_focal(179)
    assert SF.parse("{}").lens_x == SF.lens_focal(64.0)   # no camera: SimpleStochasticLense(64)


QUAD = "v 0 0 0\nv 2 0 0\nv 2 0 2\nv 0 0 2\ng quad\nf 1 2 3 4\n"
TETRA = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\ng t\nf 1 3 2\nf 1 2 4\nf 1 4 3\nf 2 3 4\n"


def test_oracle_obj_quad_fan_mirror_and_vertex_normals():
    meshes, names = SF.read_obj(QUAD)
    assert names == ["quad"] and len(meshes) == 1
    tris = SF.bake(meshes[0])
    assert len(tris) == 2
    pos = [[c[0] for c in t] for t in tris]
    # fan (0,1,2), (0,2,3) of the x-mirrored vertices
    assert [tuple(float(x) for x in p) for p in pos[0]] == [(-0.0, 0, 0), (-2, 0, 0), (-2, 0, 2)]
    assert [tuple(float(x) for x in p) for p in pos[1]] == [(-0.0, 0, 0), (-2, 0, 2), (-0.0, 0, 2)]
    # planar quad: every vertex normal is the unit plane normal, here (prev - cur) x (next - cur)
    # at corner 0 of the mirrored quad = (0,0,2) x (-2,0,0) = (0,-4,0)
    for t in tris:
        for c in t:
            np.testing.assert_allclose(np.array(c[1], np.float64), (0, -1, 0), atol=1e-7)


def test_oracle_obj_missing_normals_share_the_first_corners_vertex_normal():
    """Without vn every corner has normal index -1, and the per-object index map gives all of
    them the vertex normal of the object's first corner (vertex 1 here)."""
    meshes, _ = SF.read_obj(TETRA)
    # vertex 1 of the mirrored tetrahedron: corner cross products of faces (1 3 2), (1 2 4), (1 4 3)
    v = np.array([[0, 0, 0], [-1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    s = np.zeros(3)
    for f in ((0, 2, 1), (0, 1, 3), (0, 3, 2)):
        a, b, c = (v[i] for i in f)
        s += np.cross(c - a, b - a)        # (prev - cur) x (next - cur) at corner 0
    n0 = s / np.linalg.norm(s)
    for face in meshes[0]:
        for _, n, _t in face:
            np.testing.assert_allclose(np.array(n, np.float64), n0, atol=1e-7)


def test_oracle_obj_explicit_normals_and_polygons():
    text = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0.5 1.5 0\nv 0 1 0\nvn 0 0 1\nvn 0.6 0 0.8\ng pent\nf 1//1 2//2 3//1 4//2 5//1\n"
    meshes, names = SF.read_obj(text)
    tris = SF.bake(meshes[0])
    assert len(tris) == 3
    assert [float(x) for x in tris[0][1][1]] == [-0.6, 0.0, 0.8] or np.allclose(
        [float(x) for x in tris[0][1][1]], [-0.6, 0.0, 0.8])
    assert tris[2][2][0] == (f32(-0.0), f32(1), f32(0))


def test_oracle_obj_group_names():
    # an empty group is dropped together with every earlier name
    m, n = SF.read_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\ng A\ng B\nf 1 2 3\n")
    assert n == ["B"] and len(m) == 1
    m, n = SF.read_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\ng A\nf 1 2 3\ng B\ng C\nf 1 2 3\n")
    assert len(m) == 2 and n == ["C"]
    # faces before the first group: one more mesh than names
    m, n = SF.read_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\ng A\nf 1 2 3\n")
    assert len(m) == 2 and n == ["A"]
    # comments, smoothing groups and materials are skipped; trailing spaces trimmed from names
    m, n = SF.read_obj("# c\nv 0 0 0 # x\nv 1 0 0\nv 0 1 0\ng  named thing  \ns 1\nusemtl m\nf 1 2 3\n")
    assert n == ["named thing"] and len(SF.bake(m[0])) == 1


def test_oracle_obj_index_errors():
    with pytest.raises(SF.SceneError):
        SF.read_obj("v 0 0 0\nv 1 0 0\ng a\nf 1 2 3\n")
    with pytest.raises(SF.SceneError):
        SF.read_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\ng a\nf 1/1 2/1 3/1\n")      # no vt
    with pytest.raises(SF.SceneError):
        SF.read_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\ng a\nf 1//4 2//4 3//4\nvn 0 0 1\n")


def test_oracle_materials_and_lights_closed_forms():
    sf = SF.parse(textwrap.dedent("""
    { "materials": [
        {"name": "g", "type": "dumb_based", "preset": "glossy"},
        {"name": "gold", "type": "dumb_based", "preset": "gold", "specular": 8},
        {"type": "dumb_based", "diffuse": [0.5, 0.25, 1], "metal": 2},
        {"type": "dumb_based", "diffuse": [0.5, 0.25, 1], "metal": 0.25},
        {"type": "simple_stochastic", "diffuse": 0.8, "smoothness": 0.6, "color": [1, 0, 1]},
        {"type": "fresnel_color", "engine_color": "copper"} ],
      "lights": [
        {"type": "simple_soft_directional"},
        {"type": "simple_soft_directional", "direction": [0, -3, 4], "distance": 10, "softness": 0.5},
        {"type": "simple_spherical", "position": [1, 2, 3], "radius": 0.25, "color": [4, 5, 6]} ] }"""))
    m = sf.materials
    assert np.allclose(m[0]["diffuse"], [0.24] * 3, rtol=1e-6) and m[0]["p"][:2] == (4096.0, 0.25)
    assert m[1]["diffuse"] == (0.0, 0.0, 0.0) and m[1]["p"][:2] == (8.0, 1.0)          # metal 1: no diffuse
    assert m[2]["p"][1] == 1.0 and m[2]["diffuse"] == (0.0, 0.0, 0.0)                  # metal clamped to 1
    assert m[3]["diffuse"] == tuple(float(f32(c) * f32(0.75)) for c in (0.5, 0.25, 1))
    assert abs(m[4]["p"][0] + m[4]["p"][1] - 1.0) < 1e-6 and abs(m[4]["p"][0] / m[4]["p"][1] - 0.8 / 0.6) < 1e-6
    assert m[5]["albedo"] == tuple(float(f32(c)) for c in (0.95, 0.64, 0.54))
    L = sf.lights
    assert np.allclose(L[0]["edge_u"], np.array([1, -1, 1]) / math.sqrt(3), atol=1e-7)
    assert L[0]["dist"] == 1024.0 and L[0]["radius"] == 128.0
    assert np.allclose(L[1]["edge_u"], [0, -0.6, 0.8], atol=1e-7) and L[1]["radius"] == 5.0
    assert L[2]["corner"] == (1.0, 2.0, 3.0) and L[2]["radius"] == 0.25 and L[2]["emission"] == (4.0, 5.0, 6.0)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(textwrap.dedent(text))
    return str(p)


def test_oracle_groups_translate_and_pick_materials(tmp_path):
    """Translations compose item first, then sub-group, then group; override beats the
    item's material, fallback fills only a missing one."""
    _write(tmp_path, "m.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\ng tri\nf 1 2 3\n")
    sf = SF.load(_write(tmp_path, "s.dumb", """
    { "materials": [ {"name": "a", "type": "simple_stochastic"}, {"name": "b", "type": "fresnel_color"},
                     {"name": "c", "type": "simple_stochastic", "diffuse": 1} ],
      "groups": [ {"name": "G", "elements": [ {"mesh": {"obj": "m.obj", "object": "tri"},
                                                "transform": {"position": [0.5, 0, 0]}},
                                               {"mesh": {"obj": "m.obj", "object": "tri"}, "material": "a"} ],
                   "transform": {"position": [0, 0.25, 0]}, "fallback_material": "b"} ],
      "objects": [ "G",
                   {"group": "G", "transform": {"position": [10, 0, 0]}, "override_material": "c"} ] }"""))
    assert sf.mat_id.tolist() == [1, 0, 2, 2]
    base = np.array([[-0.0, 0, 0, -1, 0, 0, -0.0, 1, 0]], np.float32)
    shift = lambda t: np.tile(np.array(t, np.float32), 3)
    exp0 = (base + shift((0.5, 0, 0))) + shift((0, 0.25, 0))
    exp1 = base + shift((0, 0.25, 0))
    np.testing.assert_array_equal(sf.tris[0], exp0[0])
    np.testing.assert_array_equal(sf.tris[1], exp1[0])
    np.testing.assert_array_equal(sf.tris[2], (exp0 + shift((10, 0, 0)))[0])
    np.testing.assert_array_equal(sf.tris[3], (exp1 + shift((10, 0, 0)))[0])


# ----------------------------------------------------------------------------- library vs oracle
def _compare(ctx, sf):
    assert ctx.tris.shape == sf.tris.shape
    assert same(ctx.tris, sf.tris)
    assert same(ctx.normals, sf.normals)
    assert np.array_equal(ctx.mat_id, sf.mat_id)
    assert same(ctx.camera, sf.cam)
    assert same(ctx.material_table, PL.material_table(sf.materials)[: len(sf.materials)])
    assert same(ctx.light_table, PL.light_table(sf.lights)[: len(sf.lights)])
    assert ctx.max_bounces == sf.max_bounces and ctx.samples_per_pixel == sf.fsaa
    assert ctx.boxing == ["height", "width", "min", "max"][sf.boxing] and ctx.ignore_backfaces == sf.ignore_backfaces
    assert ctx.lens == sf.lens and same(ctx.lens_params, np.array(sf.lens_params, np.float32))
    assert same(ctx.cam_rows, sf.cam_rows)


def test_example_scene_fixture_is_the_reference_file():
    if not os.path.isdir(REF_EXAMPLE):
        pytest.skip("reference tree not present")
    for rel in ("0.dumb", os.path.join("files", "cargo.obj")):
        with open(os.path.join(REF_EXAMPLE, rel), "rb") as a, open(os.path.join(os.path.dirname(EXAMPLE), rel), "rb") as b:
            assert a.read() == b.read()


def test_example_scene_matches_oracle():
    ctx = D.DumbRenderContext.fromFile(EXAMPLE)
    sf = SF.load(EXAMPLE)
    _compare(ctx, sf)
    # 10 objects of cargo.obj, Box002 twice; every face a fan
    assert ctx.triangle_count == 1670 and len(sf.materials) == 9 and len(sf.lights) == 6
    assert set(ctx.mat_id.tolist()) == set(range(9))
    assert ctx.max_bounces == 2 and ctx.samples_per_pixel == (1, 1) and ctx.boxing == "height"


FULL_SCENE = """
{
  "include": ["inc/extra.dumb"],
  "materials": [
    {"name": "glossy", "type": "dumb_based", "preset": "glossy", "diffuse": [0.25, 0.1, 0.1], "fresnsel": [1, 1, 1]},
    {"name": "rough", "type": "dumb_based", "preset": "rough_gold", "metal": 0.3, "normal_color": [1, 1, 1],
     "diffuse_texture": null, "alpha_cutout_threshold": 0.25},
    {"name": "stoch", "type": "simple_stochastic", "color": [0.3, 0.6, 0.9], "shininess": 17},
    {"name": "fres", "type": "fresnel_color", "color": [0.9, 0.8, 0.7]}
  ],
  "lights": [
    {"type": "simple_soft_directional", "direction": [-1, -2, 0.5], "color": [0.5, 0.25, 0.125], "softness": 0.3},
    {"type": "simple_spherical", "color": [100, 90, 80], "position": [-3, 9, 2.5], "radius": 1.25},
    {"type": "spotlight", "color": [1, 0.9, 0.8], "luminosity": 40, "position": [2, 6, -1], "target": [0, 0, 0],
     "hotspot_angle": 25, "falloff_angle": 50, "falloff_power": 1.5, "emitter_size": 0.4, "emitter_shape": "sphere",
     "cast_shadows": false, "sample_count": 3},
    {"type": "spotlight", "direction": [0, -1, 0.2], "falloff_angle": -10, "emitter_shape": "oval", "sample_count": 99}
  ],
  "groups": [
    {"name": "pair", "elements": [
        {"mesh": {"obj": "mesh.obj", "object": "A"}, "transform": {"rotation": [10, 20, 30], "scale": [1, 2, 0.5]}},
        {"mesh": {"obj": "mesh.obj", "object": "B"}, "material": "stoch"},
        {"elements": [{"mesh": {"obj": "mesh.obj", "object": "A"}, "material": {"type": "fresnel_color"}}],
         "transform": {"position": [0, 3, 0], "rotation": [0, 45, 0]}}
     ], "fallback_material": "glossy", "transform": {"relative_position": [1, 1, 1], "rotation": [5, 0, -7]}}
  ],
  "objects": [
    {"mesh": {"obj": "mesh.obj"}, "material": "rough", "transform": {"position": [4, 0, 0]}},
    {"mesh": {"obj": "mesh.obj", "object": "B"}, "material": {"name": "inline", "type": "dumb_based", "specular": 64},
     "transform": {"position": [1, 2, 3], "rotation": [90, 0, 45], "scale": [3, 3, 3], "relative_position": [0, 0, -2]}},
    {"group": "pair", "override_material": "inline", "transform": {"position": [-5, 0, 5]}},
    "pair",
    {"elements": ["pair"], "fallback_material": "fres", "transform": {"scale": [0.5, 0.5, 0.5]}}
  ],
  "camera": {"lense": {"type": "simple_stochastic", "angle": 75},
             "transform": {"position": [1, 5, -20], "rotation": [12, -8, 3], "relative_position": [0, 0, -4]}},
  "renderer": {"resources": {"cpu": 4, "gpu": 1}, "blocks": {"cpu_cut": 2, "gpu_cut": 8},
               "pixel": {"boxing": "width", "bounces": 3.7, "samples_per_pixel": [2, 3], "ignore_backfaces": false}}
}
"""
MESH = """# two objects, one without normals (vertex normals), one with
v 0 0 0
v 1 0 0.2
v 1.3 1 0
v 0 1 -0.4
v 0.5 0.5 1
vn 0 0 1
vn 0.6 0.8 0
vt 0 0
g A
f 1 2 3 4
f 1 2 5
f 2 3 5
g B
f 3/1/1 4/1/2 5/1/1
f 4//2 1//2 5//1
"""
EXTRA = """{ "materials": [ {"name": "extra", "type": "simple_stochastic", "diffuse": 0.9, "smoothness": 0.4} ],
            "lights": [ {"type": "simple_spherical"} ] }"""


def test_full_scene_matches_oracle(tmp_path):
    path = _write(tmp_path, "scene.dumb", FULL_SCENE)
    _write(tmp_path, "mesh.obj", MESH)
    _write(tmp_path, "inc/extra.dumb", EXTRA)
    ctx = D.DumbRenderContext.fromFile(path)
    sf = SF.load(path)
    # the object without a mesh name takes every mesh of the file in the library's dictionary
    # order (std::unordered_map, as in the reference): compare that object as a set of triangles
    n_all = sum(len(SF.bake(m)) for m in sf.obj_files["obj::mesh.obj"].values())
    key = lambda a: sorted(map(bytes, np.ascontiguousarray(a).view(np.uint8).reshape(a.shape[0], -1)))
    assert key(ctx.tris[:n_all]) == key(sf.tris[:n_all])
    assert key(np.hstack([ctx.tris[:n_all], ctx.normals[:n_all]])) == key(np.hstack([sf.tris[:n_all], sf.normals[:n_all]]))
    # everything after it in push order
    rest = D.DumbRenderContext.fromText(FULL_SCENE.replace('{"mesh": {"obj": "mesh.obj"}, "material": "rough", '
                                                           '"transform": {"position": [4, 0, 0]}},', ""),
                                        str(tmp_path) + "/")
    sf_rest = SF.parse(FULL_SCENE.replace('{"mesh": {"obj": "mesh.obj"}, "material": "rough", '
                                          '"transform": {"position": [4, 0, 0]}},', ""), str(tmp_path) + "/")
    _compare(rest, sf_rest)
    assert same(ctx.tris[n_all:], sf.tris[n_all:]) and np.array_equal(ctx.mat_id, sf.mat_id)
    assert ctx.boxing == "width" and ctx.max_bounces == 3 and ctx.samples_per_pixel == (2, 3)
    assert not ctx.ignore_backfaces and ctx.light_table.shape[0] == 5 and ctx.material_table.shape[0] == 7


BAD_SCENES = [
    '{"materials": [{"type": "dumb_based", "preset": "shiny"}]}',
    '{"materials": [{"type": "unknown"}]}',
    '{"materials": [{"type": "dumb_based", "diffuse_texture": "wood"}]}',
    '{"materials": [{"type": "dumb_based", "specular": "high"}]}',
    '{"lights": [{"type": "spotlight", "luminosity": "bright"}]}',
    '{"lights": [{"type": "spotlight", "cast_shadows": 1}]}',
    '{"lights": [{"type": "simple_spherical", "radius": [1]}]}',
    '{"objects": [{"mesh": {"obj": "mesh.obj", "object": "A"}}]}',                         # no material
    '{"objects": [{"mesh": {"obj": "mesh.obj", "object": "Z"}, "material": {"type": "fresnel_color"}}]}',
    '{"objects": [{"mesh": {"obj": "missing.obj"}, "material": {"type": "fresnel_color"}}]}',
    '{"objects": [{"mesh": {"primitive": "sphere"}, "material": {"type": "fresnel_color"}}]}',
    '{"objects": ["nogroup"]}',
    '{"objects": [{"elements": [{"mesh": {"obj": "mesh.obj", "object": "A"}}]}]}',          # unresolved material
    '{"camera": {"lense": {"type": "spherical_segment", "sensor": "big"}}}',
    '{"camera": {"lense": {"type": "spherical_segment", "filter": 3}}}',
    '{"camera": {"lense": {"type": "fisheye"}}}',
    '{"camera": {"lense": {"angle": 30}}}',
    '{"renderer": {"pixel": {"boxing": "diagonal"}}}',
    '{"renderer": {"pixel": {"samples_per_pixel": [1]}}}',
    '{"renderer": {"resources": {"cpu": "many"}}}',
    '{"materials": [], "materials": []}',
    '[1, 2]',
    '{"objects": [{"mesh": {"obj": "mesh.obj", "object": "A"}, "material": "nope"}]}',
]


@pytest.mark.parametrize("text", BAD_SCENES)
def test_bad_scenes_rejected_by_both(tmp_path, text):
    _write(tmp_path, "mesh.obj", MESH)
    with pytest.raises(SF.SceneError):
        SF.parse(text, str(tmp_path) + "/")
    with pytest.raises(D.DumbRayError):
        D.DumbRenderContext.fromText(text, str(tmp_path) + "/")


@pytest.mark.parametrize("tok,expected", [("2", 2), ("2.9", 2), ("-1", -1), ("1junk", 1), ("+4e0", 4)])
def test_number_tokens_like_stream_extraction(tok, expected):
    text = '{"renderer": {"pixel": {"bounces": %s}}}' % tok
    assert D.DumbRenderContext.fromText(text).max_bounces == expected == SF.parse(text).max_bounces


def test_reference_scene_with_primitives_is_rejected():
    """The reference's Scene.dumb names OBJ files that are not in the tree and a primitive
    sphere (left unimplemented by the reference): both loaders refuse it."""
    path = "/root/reference/Scene.dumb"
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    with pytest.raises(SF.SceneError):
        SF.load(path)
    with pytest.raises(D.DumbRayError):
        D.DumbRenderContext.fromFile(path)


def test_spherical_segment_lens_settings(tmp_path):
    text = ('{"camera": {"lense": {"type": "spherical_segment", "angle": 75, "distance": 40, "sensor": 0.5, '
            '"filter": [1, 0.5, 0.25]}, "transform": {"position": [1, 2, 3], "rotation": [10, 20, 30], '
            '"scale": [1, 2, 1]}}}')
    ctx = D.DumbRenderContext.fromText(text)
    sf = SF.parse(text)
    _compare(ctx, sf)
    assert ctx.lens == PL.SPHERICAL_SEGMENT_LENS
    assert ctx.lens_params.tolist() == [75.0, 75.0, 40.0, 0.5, 1.0, 0.5, 0.25, 0.0]
    # rows carry the scale, the direction axes divide it out again
    np.testing.assert_allclose(ctx.cam_rows.reshape(3, 3)[:, 1] / 2.0, ctx.camera[3:12].reshape(3, 3)[:, 1], rtol=1e-6)
    # defaults of the lens
    d = D.DumbRenderContext.fromText('{"camera": {"lense": {"type": "spherical_segment"}}}')
    assert d.lens_params.tolist() == [60.0, 60.0, 128.0, 1.0, 1.0, 1.0, 1.0, 0.0]


def test_oracle_asin(oracle):
    for x in np.linspace(0.0, 0.999999, 400):
        assert abs(oracle.asin_f(float(np.float32(x))) - math.asin(float(np.float32(x)))) < 3e-7
    assert oracle.asin_f(0.0) == 0.0 and abs(oracle.asin_f(1.0) - math.pi / 2) < 2e-7


# ----------------------------------------------------------------------------- textures
TEX_SCENE = """
{ "materials": [
    {"name": "floor", "type": "dumb_based", "preset": "matte",
     "diffuse_texture": {"png": "tex/checker.png", "filtering": "none", "name": "checker"},
     "diffuse_tiling": [4, 4, 0], "diffuse_offset": [0.1, 0.2],
     "normal_texture": {"png": "tex/bumps.png"}, "normal_color": [1, 0.8, 0.6]},
    {"name": "leaf", "type": "dumb_based", "diffuse": [0.2, 0.7, 0.2], "metal": 0.1,
     "alpha_cutout_texture": {"png": "tex/mask.png"}, "alpha_cutout_threshold": 0.4,
     "fresnel_texture": "checker", "fresnel_tiling": [2, 2]},
    {"name": "glass", "type": "fresnel_color", "color": [0.9, 0.9, 1.0], "texture": "tex/bumps.png",
     "offset": [0.5, 0.5]},
    {"name": "plain", "type": "dumb_based", "diffuse_texture": null}
  ],
  "lights": [ {"type": "simple_soft_directional", "direction": [-0.3, -1, 0.2], "color": [1, 1, 1], "softness": 0.02},
              {"type": "simple_spherical", "color": [30, 30, 30], "position": [2, 6, -2], "radius": 0.5} ],
  "objects": [ {"mesh": {"obj": "quads.obj", "object": "floor"}, "material": "floor"},
               {"mesh": {"obj": "quads.obj", "object": "leaf"}, "material": "leaf"},
               {"mesh": {"obj": "quads.obj", "object": "leaf"}, "material": "glass",
                "transform": {"position": [2.5, 0.5, 0], "rotation": [0, 0, 20]}},
               {"mesh": {"obj": "quads.obj", "object": "leaf"}, "material": "plain",
                "transform": {"position": [-2.5, 0.2, 0.5]}} ],
  "camera": {"lense": {"type": "simple_stochastic", "angle": 60},
             "transform": {"position": [0, 1, 0], "rotation": [30, 0, 0], "relative_position": [0, 0, -8]}},
  "renderer": {"pixel": {"bounces": 2, "samples_per_pixel": [2, 2], "ignore_backfaces": false}}
}
"""
QUADS = """v -4 0 -4
v 4 0 -4
v 4 0 4
v -4 0 4
v -1 0.5 -1
v 1 0.5 -1
v 1.3 2 0
v -1 2 0.2
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vt 0.1 0.05
vt 0.9 0.1
vt 0.95 0.9
vt 0.05 0.95
vn 0 1 0
g floor
f 1/1/1 4/4/1 3/3/1 2/2/1
g leaf
f 5/5 6/6 7/7 8/8
"""


def write_textured_scene(tmp_path):
    from PIL import Image

    os.makedirs(tmp_path / "tex", exist_ok=True)
    yy, xx = np.mgrid[0:8, 0:8]
    checker = np.where(((xx // 2 + yy // 2) % 2 == 0)[..., None], [230, 60, 40, 255], [40, 60, 220, 255])
    Image.fromarray(checker.astype(np.uint8), "RGBA").save(tmp_path / "tex" / "checker.png")
    rng = np.random.default_rng(4)
    bumps = np.stack([rng.integers(90, 166, (16, 16)), rng.integers(90, 166, (16, 16)),
                      rng.integers(200, 256, (16, 16))], -1).astype(np.uint8)
    Image.fromarray(bumps, "RGB").save(tmp_path / "tex" / "bumps.png")
    mask = np.zeros((8, 8, 4), np.uint8)
    mask[..., :3] = 255
    mask[..., 3] = np.where((xx + yy) % 3 == 0, 0, 255)
    Image.fromarray(mask, "RGBA").save(tmp_path / "tex" / "mask.png")
    (tmp_path / "quads.obj").write_text(QUADS)
    path = tmp_path / "tex_scene.dumb"
    path.write_text(TEX_SCENE)
    return str(path)


def test_textured_scene_matches_oracle(tmp_path):
    path = write_textured_scene(tmp_path)
    ctx = D.DumbRenderContext.fromFile(path)
    sf = SF.load(path)
    _compare(ctx, sf)
    assert same(ctx.uv, sf.uv)
    desc, texels = oracle.pack_textures(sf.textures)
    assert np.array_equal(ctx.texture_desc, desc) and same(ctx.texels, texels)
    assert same(ctx.maps, oracle.pack_maps(sf.materials))
    # three files, the named checker reused by name, bumps.png reused by file name
    assert ctx.texture_desc.tolist() == [[0, 8, 8, 0], [64, 16, 16, 1], [320, 8, 8, 1]]
    assert ctx.maps.shape[0] == 3                                         # "plain" has no texture
    assert ctx.material_table[:, 10].view(np.int32).tolist() == [1, 2, 3, 0]
    # texels are the PNG bytes / 255
    assert ctx.texels[0].tolist() == [np.float32(230) / np.float32(255), np.float32(60) / np.float32(255),
                                      np.float32(40) / np.float32(255), 1.0]
    # u, v of the floor corners (fan 1 4 3, 1 3 2)
    np.testing.assert_array_equal(ctx.uv[0], [0, 0, 0, 1, 1, 1])


def test_parse_base_dir_with_or_without_separator(tmp_path):
    """dr_scene_file_parse resolves relative OBJ and PNG names against base_dir whether or
    not it ends in a separator; the result equals loading the file itself."""
    path = write_textured_scene(tmp_path)
    text = open(path).read()
    want = D.DumbRenderContext.fromFile(path)
    for base in (str(tmp_path), str(tmp_path) + "/"):
        got = D.DumbRenderContext.fromText(text, base)
        assert same(got.tris, want.tris) and same(got.texels, want.texels) and same(got.uv, want.uv)


@pytest.mark.parametrize("entry", ['"diffuse_texture": {"png": "tex/none.png"}',
                                   '"diffuse_texture": {"png": "tex/checker.png", "filtering": "cubic"}',
                                   '"diffuse_texture": {"file": "tex/checker.png"}',
                                   '"diffuse_texture": "unknown"',
                                   '"diffuse_texture": [1, 2]'])
def test_bad_textures_rejected_by_both(tmp_path, entry):
    write_textured_scene(tmp_path)
    text = '{"materials": [{"type": "dumb_based", %s}]}' % entry
    with pytest.raises(SF.SceneError):
        SF.parse(text, str(tmp_path) + "/")
    with pytest.raises(D.DumbRayError):
        D.DumbRenderContext.fromText(text, str(tmp_path) + "/")
