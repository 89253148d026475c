"""Render plugins: materials, lights and lenses as the device tables store them.

Each constructor mirrors a plugin of the reference and does what its constructor
does, in fp32, so the table holds the same values the reference object would:

  materials   lambert / mirror / metal          (the benchmark plugins)
              dumb_based_shader + presets       DumbBasedShader.cuh:14-17, .impl.cuh:5-16
              simple_stochastic_shader          SimpleStochasticShader.impl.cuh:4-11
              fresnel_shader                    FresnelShader.impl.cuh:6-9
              default_shader                    DefaultShader.impl.cuh:5-13
  lights      area_light                        (the benchmark plugin)
              directional_light                 SimpleDirectionalLight.impl.cuh:7-11
              soft_directional_light            SimpleSoftDirectionalLight.impl.cuh:5-10
              spherical_light                   SphericalLightEmitter.impl.cuh:5-9
              spot_light                        Spotlight.impl.cuh:7-38
  lenses      STOCHASTIC_LENS (SimpleStochasticLense), PERSPECTIVE_LENS (DefaultPerspectiveLense),
              SPHERICAL_SEGMENT_LENS (spherical_segment_lens: SphericalSegmentLense.impl.cuh:6-34)

Textures: ``texture`` / ``texture_map`` / ``with_maps`` describe a ColoredTexture's texture
part (colour, diffuse, normal and alpha-cutout maps of DumbBasedShader, FresnelShader's colour);
``maps_table`` / ``texture_tables`` give the device tables.
"""
from __future__ import annotations

import numpy as np

MATERIAL_KINDS = {"lambert": 0, "mirror": 1, "metal": 2, "dumb_based": 3, "simple_stochastic": 4, "fresnel": 5,
                  "default": 6}
LIGHT_KINDS = {"quad": 0, "area": 0, "directional": 1, "soft_directional": 2, "spherical": 3, "spot": 4}
SPOT_CAST_SHADOWS, SPOT_EMIT_FROM_SPHERE = 32, 64
STOCHASTIC_LENS, PERSPECTIVE_LENS, SPHERICAL_SEGMENT_LENS = 0, 1, 2
PLUGINS_BENCHMARK, PLUGINS_REFERENCE = 0, 1

f32 = np.float32


def _rgb(c):
    return tuple(float(f32(x)) for x in c)


def _normalized(v):
    """Vector3::normalized in fp32: v * (1 / sqrt(x*x + y*y + z*z))."""
    x, y, z = (f32(a) for a in v)
    inv = f32(1.0) / np.sqrt(f32(f32(x * x + y * y) + z * z), dtype=f32)
    return (float(x * inv), float(y * inv), float(z * inv))


# ---------------------------------------------------------------- materials
def lambert(albedo):
    return {"kind": "lambert", "albedo": _rgb(albedo)}


def mirror(reflectance):
    return {"kind": "mirror", "albedo": _rgb(reflectance)}


def metal(f0):
    return {"kind": "metal", "albedo": _rgb(f0)}


def dumb_based_shader(fresnel=(0.2, 0.2, 0.2), specular=8.0, diffuse=(0.2, 0.2, 0.2), metal=0.5):
    """DumbBasedShader(fresnelFactor, cpecular, diffuse, metal): metal clamped to [0, 1],
    diffuse scaled by (1 - metal)."""
    m = max(min(f32(metal), f32(1.0)), f32(0.0))
    d = tuple(float(f32(c) * (f32(1.0) - m)) for c in diffuse)
    return {"kind": "dumb_based", "albedo": _rgb(fresnel), "diffuse": d, "p": (float(f32(specular)), float(m), 0.0)}


def rough_gold():
    return dumb_based_shader((1.022, 0.782, 0.344), 16.0, (1.0, 1.0, 1.0), 1.0)


def glossy_gold():
    return dumb_based_shader((1.022, 0.782, 0.344), 128.0, (1.0, 1.0, 1.0), 1.0)


def glossy_finish():
    return dumb_based_shader((0.175, 0.15, 0.125), 4096.0, (0.32, 0.32, 0.32), 0.25)


def matte_finish():
    return dumb_based_shader((0.175, 0.15, 0.125), 8.0, (0.24, 0.24, 0.24), 0.05)


def _diffuse_gloss(diffuse, smoothness):
    total = f32(diffuse) + f32(smoothness)
    mul = (f32(1.0) / total) if total > f32(1.0) else f32(1.0)
    return float(f32(diffuse) * mul), float(f32(smoothness) * mul)


def simple_stochastic_shader(color=(1.0, 1.0, 1.0), diffuse=0.05, smoothness=0.95, shininess=2.0):
    d, g = _diffuse_gloss(diffuse, smoothness)
    return {"kind": "simple_stochastic", "albedo": _rgb(color), "p": (d, g, float(f32(shininess)))}


def fresnel_shader(color=(1.0, 1.0, 1.0)):
    return {"kind": "fresnel", "albedo": _rgb(color)}


def default_shader(color=(1.0, 1.0, 1.0), diffuse=0.7, smoothness=0.2, shine=8.0):
    d, g = _diffuse_gloss(diffuse, smoothness)
    return {"kind": "default", "albedo": _rgb(color), "p": (d, g, float(f32(shine)))}


def is_reference_material(m) -> bool:
    return MATERIAL_KINDS[m["kind"]] >= MATERIAL_KINDS["dumb_based"]


# ---------------------------------------------------------------- lights
def area_light(corner, edge_u, edge_v, emission):
    """One-sided quad emitting on the (edge_u x edge_v) side."""
    return {"kind": "quad", "corner": tuple(corner), "edge_u": tuple(edge_u), "edge_v": tuple(edge_v),
            "emission": _rgb(emission)}


def directional_light(color, direction, distance):
    return {"kind": "directional", "emission": _rgb(color), "edge_u": _normalized(direction),
            "dist": float(f32(distance))}


def soft_directional_light(color=(1.0, 1.0, 1.0), direction=None, distance=1024.0, softness=0.125):
    if direction is None:
        direction = _normalized((1.0, -1.0, 1.0))
    return {"kind": "soft_directional", "emission": _rgb(color), "edge_u": _normalized(direction),
            "dist": float(f32(distance)), "radius": float(f32(distance) * f32(softness))}


def spherical_light(color=(1.0, 1.0, 1.0), position=(0.0, 0.0, 0.0), radius=1.0):
    return {"kind": "spherical", "emission": _rgb(color), "corner": tuple(position), "radius": float(f32(radius))}


def spot_light(color=(1.0, 1.0, 1.0), luminosity=1500.0, position=(0.0, 0.0, 0.0), direction=(0.0, 0.0, 1.0),
               inner_angle=16.0, outer_angle=64.0, falloff_power=1.0, emitter_size=0.25, emit_from_sphere=False,
               cast_shadows=True, samples=1):
    """Lights::Spotlight(shade, lum, pos, dir, innerAngle, outerAngle, falloffPower, emitSize,
    emitFromSphere, castShadows, samples) as its constructor stores it (Spotlight.impl.cuh:7-38):
    colour * luminosity, unit direction, emitter size at least 32*VECTOR_EPSILON, angles folded
    into [0, 180] with outer >= inner, their cosines (C library cosf), samples clamped to [0, 16]."""
    emission = tuple(float(f32(c) * f32(luminosity)) for c in color)
    size = f32(emitter_size)
    if size < f32(32) * f32(0.000005):
        size = f32(32) * f32(0.000005)
    inner, outer = f32(inner_angle), f32(outer_angle)
    if inner < 0:
        inner = -inner
    if inner > 180:
        inner = f32(180)
    if outer < 0:
        outer = -inner
    if outer > 180:
        outer = f32(180)
    if outer < inner:
        outer = inner
    pi = f32(3.14159265359)
    ci = _cosf((inner * pi) / f32(180))
    co = _cosf((outer * pi) / f32(180))
    n = int(samples)
    flags = (16 if n > 16 else n) if n > 0 else 0
    flags |= SPOT_CAST_SHADOWS if cast_shadows else 0
    flags |= SPOT_EMIT_FROM_SPHERE if emit_from_sphere else 0
    return {"kind": "spot", "emission": emission, "corner": tuple(float(f32(c)) for c in position),
            "edge_u": _normalized(direction), "edge_v": (float(ci), float(co), float(f32(falloff_power))),
            "radius": float(size), "flags": flags}


def _cosf(x):
    import ctypes
    import ctypes.util

    global _LIBM
    if _LIBM is None:
        _LIBM = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
        _LIBM.cosf.restype = ctypes.c_float
        _LIBM.cosf.argtypes = [ctypes.c_float]
    return f32(_LIBM.cosf(float(x)))


_LIBM = None


def spherical_segment_lens(angle=(60.0, 60.0), distance=128.0, sensor=1.0, filter=(1.0, 1.0, 1.0)):
    """Lenses::SphericalSegmentLense(angle, focalDistance, softness, sensitivity)
    (SphericalSegmentLense.impl.cuh:6-11) as dr_render_params.lens_p."""
    ax, ay = (angle, angle) if np.isscalar(angle) else angle
    return np.array([ax, ay, distance, sensor, filter[0], filter[1], filter[2], 0.0], np.float32)


def is_reference_light(l) -> bool:
    return LIGHT_KINDS[l.get("kind", "quad")] != 0


# ---------------------------------------------------------------- textures
MAP_SLOTS = {"color": 0, "diffuse": 1, "normal": 2, "alpha": 3}


def texture(rgba, filtering="bilinear"):
    """A Texture (Texture.cuh:13-60): rgba = (H, W, 4) float32 texels, row 0 at the top, or
    (H, W, 4) uint8 bytes converted as Images::getTexturePNG does (byte / 255.0f, Images.cu:45-53)."""
    a = np.asarray(rgba)
    if a.dtype == np.uint8:
        a = a.astype(np.float32) / f32(255.0)
    a = np.ascontiguousarray(a, np.float32)
    assert a.ndim == 3 and a.shape[2] == 4 and a.shape[0] > 0 and a.shape[1] > 0
    return {"rgba": a, "filter": {"none": 0, "bilinear": 1}[filtering]}


def texture_map(index, tiling=(1.0, 1.0), offset=(0.0, 0.0)):
    """The texture part of a ColoredTexture: texture `index` of the scene's texture list."""
    return {"texture": int(index), "tiling": tuple(float(f32(t)) for t in tiling),
            "offset": tuple(float(f32(o)) for o in offset)}


def with_maps(material, color=None, diffuse=None, normal=None, alpha=None, normal_color=(1.0, 1.0, 1.0),
              alpha_threshold=0.5):
    """The material with textures: color (DumbBased fresnelColor / Fresnel colour), diffuse,
    normal (normal map, channel weights normal_color) and alpha (cutout below alpha_threshold)."""
    m = dict(material)
    m["maps"] = {"color": color, "diffuse": diffuse, "normal": normal, "alpha": alpha,
                 "normal_color": _rgb(normal_color), "alpha_threshold": float(f32(alpha_threshold))}
    return m


def maps_table(materials) -> np.ndarray:
    """(n, 24) float32 rows in dr_material_maps layout, one per material with maps, in order."""
    rows = []
    for m in materials:
        mp = m.get("maps")
        if mp is None:
            continue
        r = np.zeros(24, np.float32)
        for slot, k in MAP_SLOTS.items():
            tm = mp.get(slot)
            r[5 * k] = np.array([-1 if tm is None else tm["texture"]], np.int32).view(np.float32)[0]
            r[5 * k + 1:5 * k + 3] = 1.0
            if tm is not None:
                r[5 * k + 1:5 * k + 3] = tm["tiling"]
                r[5 * k + 3:5 * k + 5] = tm["offset"]
        r[20:23] = mp["normal_color"]
        r[23] = mp["alpha_threshold"]
        rows.append(r)
    return np.stack(rows) if rows else np.zeros((0, 24), np.float32)


def texture_tables(textures):
    """(descriptors (n, 4) int32 in dr_texture layout, texels (total, 4) float32)."""
    desc = np.zeros((len(textures), 4), np.int32)
    chunks, off = [], 0
    for i, t in enumerate(textures):
        h, w = t["rgba"].shape[:2]
        desc[i] = (off, w, h, t["filter"])
        chunks.append(t["rgba"].reshape(-1, 4))
        off += w * h
    texels = np.concatenate(chunks) if chunks else np.zeros((0, 4), np.float32)
    return desc, np.ascontiguousarray(texels, np.float32)


# ---------------------------------------------------------------- device tables
def material_table(materials) -> np.ndarray:
    """(n, 12) float32 rows in dr_material layout; reserved[0] = 1 + the material's row in
    maps_table for materials with textures."""
    t = np.zeros((max(len(materials), 1), 12), np.float32)
    k = 0
    for i, m in enumerate(materials):
        t[i, 0] = np.array([MATERIAL_KINDS[m["kind"]]], np.int32).view(np.float32)[0]
        t[i, 1:4] = m.get("albedo", (0.0, 0.0, 0.0))
        t[i, 4:7] = m.get("diffuse", (0.0, 0.0, 0.0))
        t[i, 7:10] = m.get("p", (0.0, 0.0, 0.0))
        if m.get("maps") is not None:
            k += 1
            t[i, 10] = np.array([k], np.int32).view(np.float32)[0]
    return t


def light_table(lights) -> np.ndarray:
    """(n, 16) float32 rows in dr_light layout."""
    t = np.zeros((max(len(lights), 1), 16), np.float32)
    for i, l in enumerate(lights):
        t[i, 0:3] = l.get("corner", (0.0, 0.0, 0.0))
        t[i, 3:6] = l.get("edge_u", (0.0, 0.0, 0.0))
        t[i, 6:9] = l.get("edge_v", (0.0, 0.0, 0.0))
        t[i, 9:12] = l.get("emission", (0.0, 0.0, 0.0))
        t[i, 12] = np.array([LIGHT_KINDS[l.get("kind", "quad")]], np.int32).view(np.float32)[0]
        t[i, 13] = l.get("dist", 0.0)
        t[i, 14] = l.get("radius", 0.0)
        t[i, 15] = np.array([int(l.get("flags", 0))], np.int32).view(np.float32)[0]
    return t


def plugin_set(materials, lights) -> int:
    """The plugin set a scene needs: the reference estimator once any reference plugin is used."""
    if any(is_reference_material(m) for m in materials) or any(is_reference_light(l) for l in lights):
        return PLUGINS_REFERENCE
    return PLUGINS_BENCHMARK
