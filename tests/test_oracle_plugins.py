"""The oracle's reference plugins pinned by closed forms, special cases and invariants.

The oracle (oracle/dr_render_ref.inc) restates DumbBasedShader, SimpleStochasticShader,
FresnelShader, DefaultShaderGeneric, the three reference lights and the reference
estimator with photon transparency (DumbRenderer.cu:256-443).  The reference ships no
golden vectors for any of them, so each test fixes a case whose value follows from the
mathematics alone: normal incidence, mirror configurations, pass-through photons,
sampling invariants, and whole renders of flat scenes whose pixels have closed forms.
"""
import math

import numpy as np
import pytest

import dumbray_amd as D
from dumbray_amd import plugins as PL
from dumbray_amd import scenes as S

PI_REF = np.float32(3.14159265359)

# a triangle in the plane y = 0 whose Triangle::normal (c - a) x (b - a) is +y
TRI_UP = np.array([-1.0, 0.0, -1.0, 2.0, 0.0, -1.0, -1.0, 0.0, 2.0], np.float32)
P0 = (0.0, 0.0, 0.0)
WHITE = (1.0, 1.0, 1.0)


def test_triangle_normal_convention():
    a, b, c = TRI_UP.reshape(3, 3)
    assert np.allclose(np.cross(c - a, b - a) / np.linalg.norm(np.cross(c - a, b - a)), (0, 1, 0))


# ---------------------------------------------------------------- pow and the sphere sampler
def test_pow_against_libm(oracle):
    xs = np.concatenate([np.geomspace(1e-6, 4.0, 97), [0.25, 0.5, 0.999, 1.001]])
    for y in (0.5, 2.0, 3.0, 8.0, 16.0, 128.0, 1.0 / 16.0, 1.0 / 4096.0, -1.5):
        for x in xs:
            e = float(x) ** y
            if not (1e-30 < e < 1e30) or abs(y * math.log2(x)) > 60:
                continue
            r = oracle.pow_f(float(np.float32(x)), y)
            ref = float(np.float32(x)) ** float(np.float32(y))
            assert abs(r - ref) <= 4e-6 * abs(ref), (x, y, r, ref)


def test_pow_special_cases(oracle):
    assert oracle.pow_f(0.0, 3.0) == 0.0
    assert oracle.pow_f(0.0, 0.0) == 1.0
    assert oracle.pow_f(5.0, 0.0) == 1.0
    assert oracle.pow_f(1.0, float("nan")) == 1.0
    assert math.isinf(oracle.pow_f(0.0, -2.0))
    assert abs(oracle.pow_f(-2.0, 3.0) + 8.0) < 1e-5          # odd integer power keeps the sign
    assert abs(oracle.pow_f(-0.5, 8.0) - 0.00390625) < 1e-9    # even integer power of a negative base
    assert math.isnan(oracle.pow_f(-2.0, 0.5))
    assert oracle.pow_f(1e-20, 4096.0) == 0.0                    # underflow


def test_point_on_sphere_radius_and_moments(oracle):
    g = (np.arange(64) + 0.5) / 64.0
    pts = np.array([oracle.point_on_sphere(a, b, 2.5) for a in g for b in g], np.float64)
    assert np.allclose(np.linalg.norm(pts, axis=1), 2.5, rtol=1e-6)
    u = pts / 2.5
    assert np.all(np.abs(u.mean(0)) < 2e-3)                          # centred
    assert np.allclose((u * u).mean(0), 1.0 / 3.0, atol=2e-3)       # isotropic second moments


# ---------------------------------------------------------------- constructors
def test_constructors_follow_the_reference():
    f32 = np.float32
    m = PL.simple_stochastic_shader((0.5, 0.6, 0.7), diffuse=0.8, smoothness=0.6, shininess=4.0)
    tot = f32(0.8) + f32(0.6)
    assert m["p"][0] == float(f32(0.8) * (f32(1) / tot)) and m["p"][1] == float(f32(0.6) * (f32(1) / tot))
    m = PL.default_shader(diffuse=0.3, smoothness=0.2)          # sum below 1: unchanged
    assert m["p"][:2] == (float(f32(0.3)), float(f32(0.2)))
    m = PL.dumb_based_shader((0.1, 0.2, 0.3), 32.0, (0.5, 0.5, 0.5), metal=1.7)   # metal clamped to 1
    assert m["p"][1] == 1.0 and m["diffuse"] == (0.0, 0.0, 0.0)
    m = PL.dumb_based_shader(diffuse=(0.5, 0.4, 0.2), metal=0.25)
    assert m["diffuse"] == tuple(float(f32(c) * f32(0.75)) for c in (0.5, 0.4, 0.2))
    gg = PL.glossy_gold()
    assert gg["albedo"] == tuple(float(f32(c)) for c in (1.022, 0.782, 0.344)) and gg["p"][0] == 128.0
    dl = PL.directional_light((1, 1, 1), (0.0, -3.0, 4.0), 100.0)
    assert np.allclose(dl["edge_u"], (0.0, -0.6, 0.8), atol=1e-7)
    sl = PL.soft_directional_light(distance=1024.0, softness=0.125)
    assert sl["radius"] == 128.0 and np.allclose(sl["edge_u"], np.array([1, -1, 1]) / math.sqrt(3), atol=1e-7)


# ---------------------------------------------------------------- getReflectedColor
def test_dumb_based_direct_at_normal_incidence(oracle):
    """Photon straight down, observer straight up: wh = n, fresnel = F, pow(1, s) = 1, G = 1,
    so brdf = clamp(F (s + 2) / (8 pi)) and colour = clamp(brdf metal + diffuse) * photon."""
    for spec, metal in ((8.0, 0.5), (128.0, 0.25), (2.0, 0.9)):
        F, diff, pc = (0.3, 0.5, 0.9), (0.6, 0.4, 0.2), (2.0, 1.0, 0.5)
        m = PL.dumb_based_shader(F, spec, diff, metal)
        out = oracle.reflected_color(m, TRI_UP, None, P0, (0, -1, 0), pc, (0, 1, 0), 0)
        mult = (spec + 2.0) / (2.0 * float(PI_REF)) / 4.0
        for c in range(3):
            brdf = min(max(F[c] * mult, 0.0), 1.0)
            want = min(max(brdf * metal + m["diffuse"][c], 0.0), 1.0) * pc[c]
            assert abs(out[c] - want) <= 1e-5 * max(1.0, want), (spec, metal, c, out, want)


def test_dumb_based_indirect_cases(oracle):
    m = PL.dumb_based_shader((0.3, 0.5, 0.9), 16.0, (0.6, 0.4, 0.2), 0.5)
    pc = (0.7, 0.8, 0.9)
    # a photon arriving from behind the geometric normal passes straight through
    assert np.array_equal(oracle.reflected_color(m, TRI_UP, None, P0, (0, 1, 0), pc, (0, 1, 0), 1, 2),
                          np.float32(pc))
    # observer behind the interpolated normal: black
    assert np.all(oracle.reflected_color(m, TRI_UP, None, P0, (0, -1, 0), pc, (0, -1, 0), 0) == 0)
    # diffuse sample (type 2): diffuse * cos(theta) / (1 - metal) * photon
    th = 0.6
    pd = (-math.sin(th), -math.cos(th), 0.0)
    out = oracle.reflected_color(m, TRI_UP, None, P0, pd, pc, (0, 1, 0), 1, 2)
    for c in range(3):
        want = m["diffuse"][c] * math.cos(th) / (1.0 - 0.5) * pc[c]
        assert abs(out[c] - want) < 2e-6, (out, want)
    # specular sample (type 1) seen along the reversed photon: wh = wi, fresnel = F, no clamp reached
    out = oracle.reflected_color(m, TRI_UP, None, P0, pd, pc, (math.sin(th), math.cos(th), 0.0), 1, 1)
    assert np.allclose(out, np.array((0.3, 0.5, 0.9)) * pc, rtol=1e-6)


def test_simple_stochastic_and_default_shader(oracle):
    m = PL.simple_stochastic_shader((0.5, 0.25, 1.0), diffuse=0.3, smoothness=0.6, shininess=4.0)
    d = PL.default_shader((0.5, 0.25, 1.0), diffuse=0.3, smoothness=0.6, shine=4.0)
    pc = (2.0, 2.0, 2.0)
    # photon straight down, observer along the mirror direction (up): both cosines are 1
    out = oracle.reflected_color(m, TRI_UP, None, P0, (0, -1, 0), pc, (0, 1, 0), 0)
    assert np.allclose(out, np.array((0.5, 0.25, 1.0)) * 2.0 * (0.3 + 0.6), rtol=1e-6)
    # observer perpendicular to the mirror direction: only the diffuse term
    out = oracle.reflected_color(m, TRI_UP, None, P0, (0, -1, 0), pc, (1, 0, 0), 0)
    assert np.allclose(out, np.array((0.5, 0.25, 1.0)) * 2.0 * 0.3, rtol=1e-6)
    # oblique photon at 45 degrees, observer at angle a from the mirror direction: pow(cos a, shine)
    a = 0.4
    r = np.array([math.sqrt(0.5), math.sqrt(0.5), 0.0])
    ob = (math.cos(a) * r + math.sin(a) * np.array([-math.sqrt(0.5), math.sqrt(0.5), 0.0]))
    out = oracle.reflected_color(m, TRI_UP, None, P0, (r[0], -r[1], 0.0), pc, ob, 1)
    want = np.array((0.5, 0.25, 1.0)) * 2.0 * (0.3 * math.sqrt(0.5) + 0.6 * math.cos(a) ** 4)
    assert np.allclose(out, want, rtol=1e-5)
    # DefaultShaderGeneric::getReflectedColor is the same formula (the observer point at p + ob)
    out_d = oracle.reflected_color(d, TRI_UP, None, P0, (r[0], -r[1], 0.0), pc, ob, 1)
    assert np.allclose(out_d, want, rtol=1e-5)


def test_fresnel_shader(oracle):
    m = PL.fresnel_shader((0.2, 0.4, 0.6))
    pc = (1.0, 2.0, 3.0)
    assert np.allclose(oracle.reflected_color(m, TRI_UP, None, P0, (0, -1, 0), pc, (0, 1, 0), 0),
                       np.array((0.2, 0.4, 0.6)) * pc, rtol=1e-6)       # normal incidence: F
    assert np.allclose(oracle.reflected_color(m, TRI_UP, None, P0, (-1, 0, 0), pc, (0, 1, 0), 0), pc,
                       rtol=1e-6)                                         # grazing: 1


def test_vertex_normals_are_interpolated(oracle):
    """nI = massCenter of the vertex normals at the hit point: a photon along the (tilted)
    normal of the vertex where the hit point sits sees Fresnel = F."""
    m = PL.fresnel_shader((0.25, 0.25, 0.25))
    n_tilt = np.array([0.8, 0.6, 0.0], np.float32)
    normals = np.concatenate([n_tilt, [0, 1, 0], [0, 1, 0]]).astype(np.float32)
    pa = TRI_UP[:3]
    out = oracle.reflected_color(m, TRI_UP, normals, pa, -n_tilt, WHITE, (0, 1, 0), 0)
    assert np.allclose(out, 0.25, atol=1e-5)
    out_flat = oracle.reflected_color(m, TRI_UP, None, pa, -n_tilt, WHITE, (0, 1, 0), 0)
    assert not np.allclose(out_flat, 0.25, atol=1e-3)


# ---------------------------------------------------------------- requestIndirectSamples
def _many_samples(oracle, m, d, n=400):
    out = []
    for pix in range(n):
        r = oracle.indirect_sample(m, TRI_UP, None, P0, d, pix=pix, samp=3, layer=1)
        if r is not None:
            out.append(r)
    return out


def test_dumb_based_samples(oracle):
    d = (0.3, -1.0, 0.2)
    diffuse = _many_samples(oracle, PL.dumb_based_shader(metal=0.0), d)
    assert len(diffuse) == 400 and all(t == 2 for _, t in diffuse)
    v = np.array([s for s, _ in diffuse])
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6) and np.all(v[:, 1] >= 0)
    assert abs(v[:, 1].mean() - 0.5) < 0.04                     # uniform hemisphere: E[cos] = 1/2
    spec = _many_samples(oracle, PL.dumb_based_shader(specular=4096.0, metal=1.0), d)
    assert len(spec) == 400 and all(t == 1 for _, t in spec)
    dn = np.array(d) / np.linalg.norm(d)
    mirror = dn - 2 * dn.dot((0, 1, 0)) * np.array([0, 1, 0])
    v = np.array([s for s, _ in spec])
    ang = np.arccos(np.clip(v @ mirror / np.linalg.norm(v, axis=1), -1, 1))
    assert np.all(ang < 0.1) and np.all(v[:, 1] >= 0)          # a tight lobe about the mirror direction
    # a ray that arrives from behind the interpolated normal requests no specular sample
    assert all(oracle.indirect_sample(PL.dumb_based_shader(metal=1.0), TRI_UP, None, P0, (0, 1, 0), pix=k) is None
               for k in range(20))


def test_simple_stochastic_default_fresnel_samples(oracle):
    d = (0.5, -1.0, 0.0)
    mirror = np.array([0.5, 1.0, 0.0])
    gl = _many_samples(oracle, PL.simple_stochastic_shader(diffuse=0.0, smoothness=1.0), d, 20)
    for s, t in gl:
        assert t == 0 and np.allclose(s, mirror / np.linalg.norm(mirror), atol=1e-6)
    df = _many_samples(oracle, PL.simple_stochastic_shader(diffuse=1.0, smoothness=0.0), d)
    v = np.array([s for s, _ in df])
    assert np.all(v[:, 1] >= 0) and abs(v[:, 1].mean() - 0.5) < 0.04
    ds = oracle.indirect_sample(PL.default_shader(), TRI_UP, None, P0, d)
    assert ds is not None and np.allclose(ds[0], mirror, atol=1e-6)       # Vector3::reflection, not normalized
    assert oracle.indirect_sample(PL.fresnel_shader(), TRI_UP, None, P0, d) is None


# ---------------------------------------------------------------- lights
def test_light_photons(oracle):
    p = (1.0, 2.0, 3.0)
    o, d, c = oracle.light_photon(PL.directional_light((1, 2, 3), (0, -1, 0), 100.0), p)
    assert np.allclose(o, (1.0, 102.0, 3.0)) and np.allclose(d, (0, -1, 0)) and np.allclose(c, (1, 2, 3))
    base = np.array(p) - np.array(PL.soft_directional_light()["edge_u"]) * 1024.0
    for pix in range(50):
        o, d, c = oracle.light_photon(PL.soft_directional_light(), p, pix=pix)
        assert np.linalg.norm(o - base) <= 128.0 * (1 + 1e-6)
        assert np.allclose(d, np.array(p) - o, atol=1e-3)
    for pix in range(50):
        o, d, c = oracle.light_photon(PL.spherical_light((4, 4, 4), (0, 5, 0), 0.5), p, pix=pix)
        assert np.linalg.norm(o - np.array([0, 5, 0])) <= 0.5 * (1 + 1e-6)
        delta = np.array(p) - o
        assert np.allclose(d, delta / np.linalg.norm(delta), atol=1e-6)
        assert np.allclose(c, 4.0 / delta.dot(delta), rtol=1e-5)


# ---------------------------------------------------------------- whole renders with closed forms
def _floor(y=0.0, half=20.0):
    # two triangles whose Triangle::normal is +y
    a, b, c, d = (-half, y, -half), (-half, y, half), (half, y, half), (half, y, -half)
    return np.array([a + d + b, b + d + c], np.float32)


def _camera_down(height):
    cam = S.Camera.look_at((0.0, height, 0.0), (0.0, 0.0, 0.001), fov_deg=60.0)
    return cam


def _render_ref(oracle, tris, mat_id, materials, lights, cam, W, H, **kw):
    t = oracle.Octree(tris, 16, 12)
    R = oracle.Renderer(t, mat_id, materials, lights, cam.as_array(), W, H, lens=1, **kw)
    return R.render_ref(nthreads=4)


def _camera_dirs(cam, W, H):
    """Unit camera ray directions of the perspective lens at the pixel centres (1 spp)."""
    c = cam.as_array().astype(np.float64)
    P, R, U, F, f = c[0:3], c[3:6], c[6:9], c[9:12], c[12]
    ys, xs = np.mgrid[0:H, 0:W]
    sx = (xs - W / 2.0) / H + 0.5 / H
    sy = (H / 2.0 - ys) / H + 0.5 / H
    v = np.stack([sx, sy, np.full_like(sx, f)], -1)
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    return v[..., :1] * R + v[..., 1:2] * U + v[..., 2:3] * F


def test_directional_light_on_dumb_based_floor_closed_form(oracle):
    """One layer, light straight down: every pixel is clamp(brdf(metal) + diffuse) * E with
    the observer = the reversed camera ray; computed per pixel in float64."""
    W, H = 24, 16
    spec, metal, F, diff, E = 8.0, 0.3, (0.9, 0.6, 0.3), (0.5, 0.5, 0.5), (1.5, 1.0, 0.5)
    m = PL.dumb_based_shader(F, spec, diff, metal)
    img = _render_ref(oracle, _floor(), [0, 0], [m], [PL.directional_light(E, (0, -1, 0), 50.0)], _camera_down(3.0),
                      W, H, max_bounces=0)
    dirs = _camera_dirs(_camera_down(3.0), W, H)
    n = np.array([0.0, 1.0, 0.0])
    wi = np.array([0.0, 1.0, 0.0])
    w0 = -dirs
    wh = w0 + wi
    wh /= np.linalg.norm(wh, axis=-1, keepdims=True)
    val = 1.0 - (wh @ wi)
    fres = np.array(F) + (1.0 - np.array(F)) * (val ** 5)[..., None]
    nwh, nw0 = wh @ n, w0 @ n
    dwh = (spec + 2.0) / (2.0 * float(PI_REF)) * nwh ** spec
    g = np.minimum(1.0, np.minimum(2 * nwh * nw0 / (w0 * wh).sum(-1), 2 * nwh * 1.0 / (w0 * wh).sum(-1)))
    mult = dwh * g / (4.0 * nw0 * 1.0)
    brdf = np.clip(fres * mult[..., None], 0, 1)
    want = np.clip(brdf * metal + np.array(m["diffuse"]), 0, 1) * np.array(E)
    assert np.allclose(img, want, rtol=2e-5, atol=1e-6), np.abs(img - want).max()


def test_photon_passes_through_a_simple_stochastic_blocker(oracle):
    """A SimpleStochasticShader sheet above the floor: the direct photon reaches the floor
    with colour E * albedo * diffuse (the sheet's getReflectedColor toward the photon
    direction), so the matte floor shows diffuse_floor * that colour (DumbRenderer.cu:396-413)."""
    W, H = 16, 12
    floor = PL.dumb_based_shader((0.0, 0.0, 0.0), 8.0, (0.8, 0.6, 0.4), 0.0)
    sheet = PL.simple_stochastic_shader((0.5, 0.7, 0.9), diffuse=0.6, smoothness=0.0, shininess=2.0)
    tris = np.concatenate([_floor(0.0), _floor(5.0)])
    E = (2.0, 2.0, 2.0)
    img = _render_ref(oracle, tris, [0, 0, 1, 1], [floor, sheet], [PL.directional_light(E, (0, -1, 0), 50.0)],
                      _camera_down(3.0), W, H, max_bounces=0, ignore_backfaces=False)
    want = np.array([0.8, 0.6, 0.4]) * np.array(E) * np.array([0.5, 0.7, 0.9]) * 0.6
    close = np.all(np.isclose(img, want, rtol=1e-5), axis=-1)
    assert close.mean() > 0.95, img.reshape(-1, 3)[~close.reshape(-1)][:4]
    # an opaque (benchmark Lambert) sheet blocks the light completely
    img2 = _render_ref(oracle, tris, [0, 0, 1, 1], [floor, PL.lambert((0.5, 0.5, 0.5))],
                       [PL.directional_light(E, (0, -1, 0), 50.0)], _camera_down(3.0), W, H, max_bounces=0,
                       ignore_backfaces=False)
    assert np.all(img2 == 0)


def test_reference_estimator_equals_benchmark_estimator_on_benchmark_plugins(oracle):
    """With only the benchmark plugins the reference estimator renders what the benchmark
    estimator renders (same samples; sums in a different order)."""
    sc = S.spheres_scene(4, seed=2)
    t = oracle.Octree(sc.tris, 16, 12)
    R = oracle.Renderer(t, sc.mat_id, sc.materials, sc.lights, sc.camera.as_array(), 40, 30, (2, 2), 3)
    a = R.render(nthreads=4)
    b = R.render_ref(nthreads=4)
    assert a.max() > 0
    assert np.allclose(a, b, rtol=2e-5, atol=1e-6)


# ---------------------------------------------------------------- Lights::Spotlight
def test_spot_constructor_follows_the_reference():
    s = PL.spot_light((1, 0.5, 0.25), 100.0, (1, 2, 3), (0, -2, 0), inner_angle=-30, outer_angle=10, emitter_size=0,
                      samples=40, cast_shadows=False, emit_from_sphere=True)
    assert s["emission"] == (100.0, 50.0, 25.0) and s["edge_u"] == (0.0, -1.0, 0.0)
    ci, co, _ = s["edge_v"]
    # inner folded to 30, outer < inner -> inner
    assert abs(ci - math.cos(math.radians(30))) < 1e-6 and co == ci
    assert s["radius"] == float(np.float32(32) * np.float32(0.000005))
    assert s["flags"] == 16 | PL.SPOT_EMIT_FROM_SPHERE
    s = PL.spot_light(outer_angle=-5.0, inner_angle=20.0)          # negative outer -> -inner -> inner
    assert s["edge_v"][1] == s["edge_v"][0]
    s = PL.spot_light(inner_angle=200.0, outer_angle=300.0, samples=0)
    assert s["edge_v"][0] == s["edge_v"][1] and abs(s["edge_v"][0] + 1.0) < 1e-6 and s["flags"] == PL.SPOT_CAST_SHADOWS


def test_spot_photon_on_axis_closed_form(oracle):
    """A point on the axis at distance D of a disc spot: colour E * s^2 / (n * (s^2 + D^2 (1 - cos outer))),
    direction the axis, no falloff inside the hotspot."""
    D_, E, size, n = 10.0, (100.0, 200.0, 300.0), 0.01, 4
    s = PL.spot_light(E, 1.0, (0, 5, 0), (0, -1, 0), 20.0, 40.0, 1.0, size, samples=n)
    base = np.float64(np.float32(size)) ** 2
    want = np.array(E) * base / (n * (base + D_ * D_ * (1 - math.cos(math.radians(40)))))
    for k in range(n):
        o, d, c = oracle.light_photon(s, (0, 5 - D_, 0), pix=3, sample=k)
        assert np.linalg.norm(o - np.array([0, 5, 0])) <= size * (1 + 1e-6)
        assert np.allclose(d, (0, -1, 0), atol=2e-3)
        assert np.allclose(c, want, rtol=1e-3)
    # distinct samples draw distinct emitter points
    pts = {tuple(oracle.light_photon(s, (0, -5, 0), pix=3, sample=k)[0]) for k in range(n)}
    assert len(pts) == n


def test_spot_photon_falloff_and_cone(oracle):
    inner, outer, power = 20.0, 40.0, 2.5
    s = PL.spot_light((1, 1, 1), 1.0, (0, 0, 0), (0, 0, 1), inner, outer, power, 1e-6)
    for ang in (10.0, 25.0, 30.0, 39.0, 45.0, 90.0, 170.0):
        p = (math.sin(math.radians(ang)) * 10, 0.0, math.cos(math.radians(ang)) * 10)
        r = oracle.light_photon(s, p)
        if ang > outer:
            assert r is None
            continue
        _, d, c = r
        cosine = math.cos(math.radians(ang))
        base = np.float64(np.float32(32) * np.float32(0.000005)) ** 2
        val = base / (base + 100.0 * (1 - math.cos(math.radians(outer))))
        if ang > inner:
            ci, co = math.cos(math.radians(inner)), math.cos(math.radians(outer))
            a = (ci - cosine) / (ci - co)
            b = 1 - a
            val *= (a * b * b + b * (1 - a * a)) ** power
        assert np.allclose(c, val, rtol=2e-4), (ang, c, val)


def test_spot_sphere_emitter(oracle):
    s = PL.spot_light((2, 2, 2), 1.0, (1, 1, 1), (1, 0, 0), 30.0, 60.0, 1.0, 0.5, emit_from_sphere=True, samples=2)
    for pix in range(40):
        o, d, c = oracle.light_photon(s, (11, 1, 1), pix=pix, sample=pix % 2)
        assert np.linalg.norm(o - 1.0) <= 0.5 * (1 + 1e-6)
        delta = np.array([11.0, 1, 1]) - o
        # sphere emitters have no disc term: E * s^2 / (n * d^2 * (1 - cos outer))
        want = 2.0 * 0.25 / (2 * delta.dot(delta) * (1 - math.cos(math.radians(60))))
        assert np.allclose(c, want, rtol=1e-4)


def test_spot_without_shadows_ignores_blockers(oracle):
    """A spot that casts no shadows lights the floor through an opaque sheet; with shadows
    the sheet blocks it; without the sheet both agree."""
    W, H = 16, 12
    floor = PL.dumb_based_shader((0.0, 0.0, 0.0), 8.0, (0.8, 0.6, 0.4), 0.0)
    tris_open = _floor(0.0)
    tris = np.concatenate([_floor(0.0), _floor(2.0, half=0.5)])
    kw = dict(max_bounces=0, ignore_backfaces=False)
    # light off to the side, so the sheet's shadow falls on floor the camera sees
    lit = lambda shadows: PL.spot_light((1, 1, 1), 50.0, (1.0, 4, 0), (0, -1, 0), 70.0, 85.0, 1.0, 0.05,
                                        cast_shadows=shadows, samples=3)
    a = _render_ref(oracle, tris_open, [0, 0], [floor], [lit(True)], _camera_down(3.0), W, H, **kw)
    b = _render_ref(oracle, tris_open, [0, 0], [floor], [lit(False)], _camera_down(3.0), W, H, **kw)
    assert a.max() > 0 and np.array_equal(a, b)
    blocker = PL.lambert((0.5, 0.5, 0.5))
    c = _render_ref(oracle, tris, [0, 0, 1, 1], [floor, blocker], [lit(True)], _camera_down(3.0), W, H, **kw)
    d = _render_ref(oracle, tris, [0, 0, 1, 1], [floor, blocker], [lit(False)], _camera_down(3.0), W, H, **kw)
    floor_px = np.all(np.isclose(d, b), axis=-1)            # pixels that see the floor, lit regardless
    assert floor_px.mean() > 0.3
    assert (c[floor_px].sum(-1) < d[floor_px].sum(-1)).any()  # some of them are shadowed by the sheet


# ---------------------------------------------------------------- Lenses::SphericalSegmentLense
def test_spherical_segment_lens_rays_closed_form(oracle):
    """Camera rays of the spherical segment lens against a float64 statement of
    SphericalSegmentLense::getPixelSamples + Ray >>= Transform: angles alpha * screen position,
    target on the focal sphere, origin on the asin(u)*sensor disc (point map by the rows),
    direction through the axes."""
    W, H, fx, fy = 40, 30, 2, 3
    cam = S.Camera.look_at((1.0, 2.0, -3.0), (0.2, 0.5, 4.0), fov_deg=50.0)
    rows = np.array([0.9, 0.1, -0.2, 0.05, 1.1, 0.3, 0.2, -0.25, 0.95], np.float32)   # any point map
    lp = PL.spherical_segment_lens((70.0, 40.0), 25.0, 0.6, (1.0, 0.5, 0.25))
    t = oracle.Octree(_floor(), 16, 12)
    R = oracle.Renderer(t, [0, 0], [PL.lambert((1, 1, 1))], [], cam.as_array(), W, H, (fx, fy), 1, lens=2,
                        lens_params=lp, cam_rows=rows)
    c = cam.as_array().astype(np.float64)
    rng = np.random.default_rng(5)
    for _ in range(60):
        px, py, i, j = rng.integers(0, W), rng.integers(0, H), rng.integers(0, fx), rng.integers(0, fy)
        u = rng.uniform(0, 1, 4).astype(np.float32)
        o, d = oracle.camera_ray(R, px, py, i, j, u)
        ps = 1.0 / H
        sx = (px - W / 2.0) * ps + (i + 0.5) * ps / fx
        sy = (H / 2.0 - py) * ps + (j + 0.5) * ps / fy
        jx, jy = (u[0] * 2 - 1) * ps / fx / 2, (u[1] * 2 - 1) * ps / fy / 2
        h, v = math.radians((sx + jx) * 70.0), math.radians((sy + jy) * 40.0)
        target = np.array([math.sin(h) * math.cos(v), math.sin(v), math.cos(h) * math.cos(v)]) * 25.0
        a = 2 * math.pi * float(u[2])
        org = np.array([math.cos(a), math.sin(a), 0.0]) * math.asin(float(u[3])) * 0.6
        dl = (target - org) / np.linalg.norm(target - org)
        want_o = rows.astype(np.float64).reshape(3, 3) @ org + c[0:3]
        want_d = dl[0] * c[3:6] + dl[1] * c[6:9] + dl[2] * c[9:12]
        assert np.allclose(o, want_o, atol=2e-5) and np.allclose(d, want_d, atol=2e-5), (o, want_o, d, want_d)


def test_spherical_segment_lens_filter_scales_the_image(oracle):
    sc = S.spheres_scene(4, seed=2)
    kw = dict(lens=2, cam_rows=None)
    t = oracle.Octree(sc.tris, 16, 12)
    mk = lambda f: oracle.Renderer(t, sc.mat_id, sc.materials, sc.lights, sc.camera.as_array(), 24, 16, (1, 1), 2,
                                   lens_params=PL.spherical_segment_lens(60.0, 50.0, 0.1, f), **kw)
    a = mk((1.0, 1.0, 1.0)).render(nthreads=4)
    b = mk((0.5, 0.25, 2.0)).render(nthreads=4)
    assert a.max() > 0
    np.testing.assert_allclose(b, a * np.array([0.5, 0.25, 2.0], np.float32), rtol=1e-6)


# ---------------------------------------------------------------- textures
def _tex(rgb_rows, filtering="none"):
    a = np.array(rgb_rows, np.float32)
    if a.shape[-1] == 3:
        a = np.concatenate([a, np.ones(a.shape[:2] + (1,), np.float32)], -1)
    return PL.texture(a, filtering)


def test_texture_sample_closed_forms(oracle):
    red, blue, green, white = (1, 0, 0), (0, 0, 1), (0, 1, 0), (1, 1, 1)
    row = _tex([[red, blue]], "none")
    d, t = oracle.pack_textures([row])
    s = lambda x, y: oracle.texture_sample(d[0], t, x, y)[:3].tolist()
    # nearest: the texel of the 2x2 cell the offset is closer to
    assert s(0.1, 0.5) == [1, 0, 0] and s(0.4, 0.5) == [0, 0, 1] and s(0.6, 0.5) == [0, 0, 1]
    assert s(0.9, 0.5) == [1, 0, 0]                         # wraps to texel 0
    assert s(1.1, 0.5) == s(0.1, 0.5) and s(-0.6, 0.5) == s(0.4, 0.5)
    bil = _tex([[red, blue]], "bilinear")
    d, t = oracle.pack_textures([bil])
    np.testing.assert_allclose(oracle.texture_sample(d[0], t, 0.25, 0.5)[:3], [0.5, 0, 0.5], atol=1e-7)
    np.testing.assert_allclose(oracle.texture_sample(d[0], t, 0.1, 0.5)[:3], [0.8, 0, 0.2], atol=1e-6)
    # y runs up: row 0 (the top of the image) is sampled near v = 1
    col = _tex([[red], [blue]], "none")
    d, t = oracle.pack_textures([col])
    assert oracle.texture_sample(d[0], t, 0.5, 0.9)[:3].tolist() == [1, 0, 0]
    assert oracle.texture_sample(d[0], t, 0.5, 0.4)[:3].tolist() == [0, 0, 1]
    # bilinear interior points against a float64 statement of the same grid
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 1, (5, 7, 4)).astype(np.float32)
    d, t = oracle.pack_textures([PL.texture(img, "bilinear")])
    for _ in range(200):
        x, y = rng.uniform(-3, 3, 2)
        fx, fy = x - math.floor(x), 1.0 - (y - math.floor(y))
        px, py = fx * 7, fy * 5
        x0, y0 = int(px) % 7, int(py) % 5
        ox, oy = px - int(px), py - int(py)
        x1, y1 = (x0 + 1) % 7, (y0 + 1) % 5
        want = (img[y0, x0] * (1 - ox) * (1 - oy) + img[y1, x0] * (1 - ox) * oy + img[y0, x1] * ox * (1 - oy)
                + img[y1, x1] * ox * oy)
        if min(ox, 1 - ox, oy, 1 - oy) < 1e-4:
            continue                                        # float rounding of the cell edge
        np.testing.assert_allclose(oracle.texture_sample(d[0], t, x, y), want, atol=2e-5)


def _tex_floor_render(oracle, material, textures=None, uv=None, W=16, H=12, extra_tris=None, extra_mats=None,
                      bounces=0):
    tris = _floor() if extra_tris is None else np.concatenate([_floor(), extra_tris])
    mats = [material] + (extra_mats or [])
    mid = [0, 0] + ([1] * (len(tris) - 2))
    t = oracle.Octree(tris, 16, 12)
    R = oracle.Renderer(t, mid, mats, [PL.directional_light((1.5, 1.0, 0.5), (0, -1, 0), 50.0)],
                        _camera_down(3.0).as_array(), W, H, (1, 1), bounces, lens=1, ignore_backfaces=False,
                        textures=textures, uv=uv)
    return R.render_ref(nthreads=4)


def _floor_uv(n=2):
    # floor vertices (-20,0,-20) (20,0,-20) (-20,0,20) / (-20,0,20) (20,0,-20) (20,0,20): u = (x+20)/40, v = (z+20)/40
    tris = _floor().reshape(-1, 3, 3)
    uv = np.stack([(tris[..., 0] + 20) / 40, (tris[..., 2] + 20) / 40], -1).reshape(-1, 6)
    return uv.astype(np.float32)


def test_constant_diffuse_texture_scales_the_diffuse_colour(oracle):
    base = PL.dumb_based_shader((0.0, 0.0, 0.0), 8.0, (0.8, 0.6, 0.4), 0.0)
    c = (0.5, 0.25, 0.75)
    tex = [_tex([[c]], "bilinear")]
    textured = PL.with_maps(base, diffuse=PL.texture_map(0, (3.0, 2.0), (0.1, 0.2)))
    a = _tex_floor_render(oracle, textured, tex, _floor_uv())
    scaled = PL.dumb_based_shader((0.0, 0.0, 0.0), 8.0,
                                  tuple(float(np.float32(x) * np.float32(y)) for x, y in zip((0.8, 0.6, 0.4), c)), 0.0)
    b = _tex_floor_render(oracle, scaled)
    assert a.max() > 0
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-7)


def test_checker_texture_follows_the_texture_coordinates(oracle):
    """A 2x2 checker on the floor with u, v from x, z: each pixel shows the checker colour of the
    floor point its camera ray meets."""
    A, B = (1.0, 0.2, 0.2), (0.2, 0.2, 1.0)
    tex = [_tex([[A, B], [B, A]], "none")]
    m = PL.with_maps(PL.dumb_based_shader((0.0, 0.0, 0.0), 8.0, (1.0, 1.0, 1.0), 0.0),
                     diffuse=PL.texture_map(0, (8.0, 8.0)))
    W, H = 24, 18
    img = _tex_floor_render(oracle, m, tex, _floor_uv(), W, H)
    dirs = _camera_dirs(_camera_down(3.0), W, H)
    hit = np.array([0.0, 3.0, 0.0]) + dirs * (3.0 / -dirs[..., 1:2])       # floor points
    u, v = (hit[..., 0] + 20) / 40 * 8, (hit[..., 2] + 20) / 40 * 8
    # nearest texel of the 2x2 texture: column from u, row from 1 - v (row 0 at the top)
    col = np.floor(((u % 1.0) * 2) + 0.5).astype(int) % 2
    row = np.floor(((1.0 - (v % 1.0)) * 2) + 0.5).astype(int) % 2
    expect_A = (col == row)
    cos = 1.0
    want = np.where(expect_A[..., None], np.array(A), np.array(B)) * cos * np.array([1.5, 1.0, 0.5])
    frac_u, frac_v = (u * 2) % 1.0, (v * 2) % 1.0
    clear = (np.abs(frac_u - 0.5) > 0.02) & (np.abs(frac_v - 0.5) > 0.02)   # away from texel edges
    assert clear.mean() > 0.5
    np.testing.assert_allclose(img[clear], np.clip(want[clear], 0, None), rtol=1e-5, atol=1e-6)


def test_alpha_cutout_sheet_lets_light_and_view_through(oracle):
    """A sheet cut out everywhere: camera rays continue through it as its indirect sample
    (type 3, colour passed on unchanged) and light photons pass it, so the floor looks as if
    the sheet were not there (one bounce: the floor seen through the sheet is one layer
    deeper and keeps its direct light)."""
    floor = PL.dumb_based_shader((0.0, 0.0, 0.0), 8.0, (0.8, 0.6, 0.4), 0.0)
    sheet = PL.dumb_based_shader((0.2, 0.2, 0.2), 8.0, (0.9, 0.9, 0.9), 0.0)
    cut = [_tex(np.zeros((1, 1, 4), np.float32), "none")]                  # alpha 0 everywhere
    sheet_tris = _floor(2.0, half=0.8)
    uv = np.zeros((4, 6), np.float32)
    alone = _tex_floor_render(oracle, floor, bounces=1)
    with_cut = _tex_floor_render(oracle, floor, cut, uv, extra_tris=sheet_tris, bounces=1,
                                 extra_mats=[PL.with_maps(sheet, alpha=PL.texture_map(0), alpha_threshold=0.5)])
    assert alone.max() > 0
    np.testing.assert_allclose(with_cut, alone, rtol=1e-5, atol=1e-7)
    solid = [_tex(np.ones((1, 1, 4), np.float32), "none")]                 # alpha 1: above the threshold
    with_solid = _tex_floor_render(oracle, floor, solid, uv, extra_tris=sheet_tris, bounces=1,
                                   extra_mats=[PL.with_maps(sheet, alpha=PL.texture_map(0), alpha_threshold=0.5)])
    assert np.abs(with_solid - alone).max() > 1e-3
    # without a bounce the cut-out sheet shows black: passing through is an indirect sample
    black = _tex_floor_render(oracle, floor, cut, uv, extra_tris=sheet_tris, bounces=0,
                              extra_mats=[PL.with_maps(sheet, alpha=PL.texture_map(0), alpha_threshold=0.5)])
    assert black.min() == 0.0 and (black.sum(-1) == 0).mean() > 0.2


def test_flat_normal_map_keeps_the_normal_and_a_tilted_one_turns_it(oracle):
    base = PL.dumb_based_shader((0.3, 0.3, 0.3), 16.0, (0.6, 0.6, 0.6), 0.4)
    plain = _tex_floor_render(oracle, base)
    flat = [_tex([[(0.5, 0.5, 1.0)]], "bilinear")]          # (r, g, b) = (0.5, 0.5, 1): weights (1, 0, 0)
    a = _tex_floor_render(oracle, PL.with_maps(base, normal=PL.texture_map(0)), flat, _floor_uv())
    np.testing.assert_allclose(a, plain, rtol=1e-6, atol=1e-7)
    tilted = [_tex([[(0.9, 0.5, 0.6)]], "bilinear")]
    b = _tex_floor_render(oracle, PL.with_maps(base, normal=PL.texture_map(0)), tilted, _floor_uv())
    assert np.abs(b - plain).max() > 1e-3
