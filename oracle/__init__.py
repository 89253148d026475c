"""ctypes binding of the CPU oracle (liboracle.so built from dr_oracle*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The dumbray_amd package
never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.ora_tri_cast.restype = ctypes.c_int
        L.ora_tri_cast.argtypes = [F32P, F32P, F32P, ctypes.c_int, ctypes.POINTER(ctypes.c_float), F32P]
        L.ora_tri_contains.restype = ctypes.c_int
        L.ora_tri_contains.argtypes = [F32P, F32P]
        L.ora_aabb_cast_preinv.restype = ctypes.c_float
        L.ora_aabb_cast_preinv.argtypes = [F32P, F32P, F32P]
        L.ora_aabb_intersects_tri.restype = ctypes.c_int
        L.ora_aabb_intersects_tri.argtypes = [F32P, F32P]
        L.ora_intersection_bounds.argtypes = [F32P, F32P, F32P]
        L.ora_intersection_center.argtypes = [F32P, F32P, F32P]
        L.ora_clip_tri.restype = ctypes.c_int
        L.ora_clip_tri.argtypes = [F32P, F32P, F32P]
        L.ora_tri_get_masses.argtypes = [F32P, F32P, F32P]
        L.ora_tri_mass_center_w.argtypes = [F32P, F32P, F32P]
        L.ora_octree_build.restype = vp
        L.ora_octree_build.argtypes = [F32P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_octree_free.argtypes = [vp]
        L.ora_octree_node_count.restype = ctypes.c_int
        L.ora_octree_node_count.argtypes = [vp]
        L.ora_octree_list_total.restype = ctypes.c_int
        L.ora_octree_list_total.argtypes = [vp]
        L.ora_octree_export.argtypes = [vp, F32P, I32P, I32P, I32P]
        L.ora_octree_cast.argtypes = [vp, F32P, ctypes.c_int, ctypes.c_int, vp, I32P, F32P, vp, vp, vp, vp,
                                      ctypes.c_int]
        L.ora_brute_cast.argtypes = [F32P, ctypes.c_int, F32P, ctypes.c_int, ctypes.c_int, vp, I32P, F32P, vp,
                                     ctypes.c_int]
        if hasattr(L, "ora_render"):
            _bind_render(L)
        _LIB = L
    return _LIB


class OraMaterial(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("albedo", ctypes.c_float * 3), ("diffuse", ctypes.c_float * 3),
                ("p", ctypes.c_float * 3), ("reserved", ctypes.c_int32 * 2)]


class OraLight(ctypes.Structure):
    _fields_ = [("corner", ctypes.c_float * 3), ("edge_u", ctypes.c_float * 3), ("edge_v", ctypes.c_float * 3),
                ("emission", ctypes.c_float * 3), ("kind", ctypes.c_int32), ("dist", ctypes.c_float),
                ("radius", ctypes.c_float), ("reserved", ctypes.c_int32)]


class OraTextures(ctypes.Structure):
    _fields_ = [("uv", ctypes.c_void_p), ("maps", ctypes.c_void_p), ("nmaps", ctypes.c_int), ("tex", ctypes.c_void_p),
                ("ntex", ctypes.c_int), ("texels", ctypes.c_void_p)]


def texture_sample(desc, texels, x, y):
    """Texture::operator()(Vector2) of texture `desc` (offset, w, h, filter) over texels (n, 4)."""
    d = np.ascontiguousarray(desc, np.int32).reshape(4)
    t = np.ascontiguousarray(texels, np.float32).reshape(-1, 4)
    out = np.zeros(4, np.float32)
    lib().ora_texture_sample(d.ctypes.data, t.ctypes.data, ctypes.c_float(x), ctypes.c_float(y), out.ctypes.data)
    return out


class OraRenderCfg(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("fsaa_x", ctypes.c_int32),
                ("fsaa_y", ctypes.c_int32), ("max_bounces", ctypes.c_int32), ("rr_start", ctypes.c_int32),
                ("rr_survive", ctypes.c_float), ("cull_backfaces", ctypes.c_int32), ("seed_lo", ctypes.c_uint32),
                ("seed_hi", ctypes.c_uint32), ("request", ctypes.c_uint32), ("iteration", ctypes.c_int32),
                ("cam", ctypes.c_float * 13), ("lens", ctypes.c_int32), ("boxing", ctypes.c_int32),
                ("lens_p", ctypes.c_float * 8), ("cam_rows", ctypes.c_float * 9)]


KINDS = {"lambert": 0, "mirror": 1, "metal": 2, "dumb_based": 3, "simple_stochastic": 4, "fresnel": 5,
         "default": 6}
LIGHT_KINDS = {"quad": 0, "area": 0, "directional": 1, "soft_directional": 2, "spherical": 3, "spot": 4}


def pack_material(m, out=None, maps_row=0):
    """OraMaterial from a material dict: kind, albedo and the optional diffuse / p fields;
    maps_row = 1 + the material's row of the maps table (0: none)."""
    r = out if out is not None else OraMaterial()
    r.kind = KINDS[m["kind"]]
    r.reserved[0] = int(maps_row)
    for c in range(3):
        r.albedo[c] = float(np.float32(m.get("albedo", (0, 0, 0))[c]))
        r.diffuse[c] = float(np.float32(m.get("diffuse", (0, 0, 0))[c]))
        r.p[c] = float(np.float32(m.get("p", (0, 0, 0))[c]))
    return r


def pack_maps(materials):
    """Rows of ora_material_maps (24 floats: 4 maps of texture, tiling, offset; normal colour;
    alpha threshold) of the materials that have "maps", in order."""
    slots = ("color", "diffuse", "normal", "alpha")
    rows = []
    for m in materials:
        mp = m.get("maps")
        if mp is None:
            continue
        r = np.zeros(24, np.float32)
        for k, name in enumerate(slots):
            tm = mp.get(name)
            r[5 * k:5 * k + 1] = np.array([-1 if tm is None else tm["texture"]], np.int32).view(np.float32)
            r[5 * k + 1:5 * k + 3] = 1.0                     # a map without texture: tiling 1, offset 0
            if tm is not None:
                r[5 * k + 1:5 * k + 5] = list(tm["tiling"]) + list(tm["offset"])
        r[20:23] = mp["normal_color"]
        r[23] = mp["alpha_threshold"]
        rows.append(r)
    return np.stack(rows) if rows else np.zeros((0, 24), np.float32)


def pack_textures(textures):
    """(descriptors (n,4) int32: texel offset, width, height, filter; texels (m,4) float32)."""
    desc, chunks, off = [], [], 0
    for t in textures:
        h, w = t["rgba"].shape[:2]
        desc.append((off, w, h, t["filter"]))
        chunks.append(np.asarray(t["rgba"], np.float32).reshape(-1, 4))
        off += w * h
    return (np.array(desc, np.int32).reshape(-1, 4),
            np.concatenate(chunks) if chunks else np.zeros((0, 4), np.float32))


def pack_light(l, out=None):
    """OraLight from a light dict: kind (default quad) and its fields."""
    r = out if out is not None else OraLight()
    r.kind = LIGHT_KINDS[l.get("kind", "quad")]
    for c in range(3):
        r.corner[c] = l.get("corner", (0, 0, 0))[c]
        r.edge_u[c] = l.get("edge_u", (0, 0, 0))[c]
        r.edge_v[c] = l.get("edge_v", (0, 0, 0))[c]
        r.emission[c] = l.get("emission", (0, 0, 0))[c]
    r.dist = float(l.get("dist", 0.0))
    r.radius = float(l.get("radius", 0.0))
    r.reserved = int(l.get("flags", 0))
    return r


def _bind_render(L):
    vp = ctypes.c_void_p
    L.ora_render.restype = ctypes.c_int
    L.ora_render.argtypes = [vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int]
    L.ora_backward.restype = ctypes.c_int
    L.ora_backward.argtypes = [vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int]
    L.ora_render_ref.restype = ctypes.c_int
    L.ora_render_ref.argtypes = [vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int]
    L.ora_render_ref_tex.restype = ctypes.c_int
    L.ora_render_ref_tex.argtypes = [vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int]
    L.ora_texture_sample.argtypes = [vp, vp, ctypes.c_float, ctypes.c_float, vp]
    L.ora_philox.argtypes = [vp, vp, vp]
    L.ora_sincos_2pi.argtypes = [ctypes.c_float, vp, vp]
    L.ora_pow.restype = ctypes.c_float
    L.ora_pow.argtypes = [ctypes.c_float, ctypes.c_float]
    L.ora_asin.restype = ctypes.c_float
    L.ora_asin.argtypes = [ctypes.c_float]
    L.ora_camera_ray.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp]
    L.ora_point_on_sphere.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, vp]
    L.ora_reflected_color.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
    L.ora_indirect_sample.restype = ctypes.c_int
    L.ora_indirect_sample.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, vp, vp]
    L.ora_light_photon.restype = ctypes.c_int
    L.ora_light_photon.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_int, vp, vp, vp]


def _flat32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))
    assert n is None or a.size == n
    return a


def pow_f(x, y):
    """The oracle's pow (the reference shaders' pow calls)."""
    return lib().ora_pow(ctypes.c_float(x), ctypes.c_float(y))


def camera_ray(renderer, px, py, i, j, u):
    """The oracle's camera ray (origin, direction) of a Renderer's settings for lens uniforms u[4]."""
    uu = _flat32(u, 4)
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().ora_camera_ray(ctypes.addressof(renderer.cfg), int(px), int(py), int(i), int(j), uu.ctypes.data,
                         o.ctypes.data, d.ctypes.data)
    return o, d


def asin_f(x):
    """The oracle's arcsin on [0, 1] (SphericalSegmentLense aperture radius)."""
    return lib().ora_asin(ctypes.c_float(x))


def point_on_sphere(ua, ub, radius=1.0):
    out = np.zeros(3, np.float32)
    lib().ora_point_on_sphere(ctypes.c_float(ua), ctypes.c_float(ub), ctypes.c_float(radius), out.ctypes.data)
    return out


def reflected_color(material, tri, normals, p, photon_dir, photon_color, observer, photon_type, sample_type=0):
    """getReflectedColor of a reference shader (photon_type 0 direct, 1 indirect); normals None = flat."""
    m = pack_material(material)
    t, pp, pd, pc, ob = _flat32(tri, 9), _flat32(p, 3), _flat32(photon_dir, 3), _flat32(photon_color, 3), _flat32(observer, 3)
    nr = None if normals is None else _flat32(normals, 9)
    out = np.zeros(3, np.float32)
    lib().ora_reflected_color(ctypes.addressof(m), t.ctypes.data, None if nr is None else nr.ctypes.data,
                              pp.ctypes.data, pd.ctypes.data, pc.ctypes.data, ob.ctypes.data, int(photon_type),
                              int(sample_type), out.ctypes.data)
    return out


def _cfg(seed=0x5EED, request=0):
    cfg = OraRenderCfg()
    cfg.seed_lo, cfg.seed_hi = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    cfg.request = request
    return cfg


def indirect_sample(material, tri, normals, p, d, pix=0, samp=0, layer=0, seed=0x5EED):
    """requestIndirectSamples of a reference shader: (direction, sample type) or None."""
    m = pack_material(material)
    t, pp, dd = _flat32(tri, 9), _flat32(p, 3), _flat32(d, 3)
    nr = None if normals is None else _flat32(normals, 9)
    cfg = _cfg(seed)
    out = np.zeros(3, np.float32)
    ty = ctypes.c_int()
    ok = lib().ora_indirect_sample(ctypes.addressof(m), t.ctypes.data, None if nr is None else nr.ctypes.data,
                                   pp.ctypes.data, dd.ctypes.data, ctypes.addressof(cfg), pix, samp, layer,
                                   out.ctypes.data, ctypes.byref(ty))
    return (out, ty.value) if ok else None


def light_photon(light, p, index=0, pix=0, samp=0, layer=0, seed=0x5EED, sample=0):
    """getVertexPhotons of a light toward p (photon `sample` of a multi-sample light):
    (origin, direction, colour) or None."""
    L = pack_light(light)
    pp = _flat32(p, 3)
    cfg = _cfg(seed)
    o, d, c = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32)
    ok = lib().ora_light_photon(ctypes.addressof(L), int(index), int(sample), pp.ctypes.data, ctypes.addressof(cfg),
                                pix, samp, layer, o.ctypes.data, d.ctypes.data, c.ctypes.data)
    return (o, d, c) if ok else None


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().ora_philox(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def sincos_2pi(u):
    s = ctypes.c_float()
    c = ctypes.c_float()
    lib().ora_sincos_2pi(ctypes.c_float(u), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


class Renderer:
    """Oracle path estimator over an oracle Octree (same settings as dumbray_amd.DumbRenderer)."""

    def __init__(self, tree, mat_id, materials, lights, cam_array, width, height, samples_per_pixel=(1, 1),
                 max_bounces=2, rr_start=1 << 30, rr_survive=1.0, ignore_backfaces=True, seed=0x5EED, lens=0,
                 normals=None, boxing=0, lens_params=None, cam_rows=None, textures=None, uv=None, maps=None):
        self.tree = tree
        self.mat_id = np.ascontiguousarray(mat_id, np.int32)
        self.mats = (OraMaterial * max(len(materials), 1))()
        k = 0
        for i, m in enumerate(materials):
            row = 0
            if m.get("maps") is not None or m.get("maps_row"):
                k += 1
                row = m.get("maps_row", k)
            pack_material(m, self.mats[i], row)
        self.nmat = len(materials)
        self.lights = (OraLight * max(len(lights), 1))()
        for i, l in enumerate(lights):
            pack_light(l, self.lights[i])
        self.nlights = len(lights)
        self.normals = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1)
        # textures: texture tables (descriptors (n,4) int32, texels (m,4) float32) or None; maps:
        # dr_material_maps rows (k,24) float32 (the material's reserved[0] = 1 + row); uv (ntris,6)
        self._tx = None
        if maps is None and any(m.get("maps") is not None for m in materials):
            maps = pack_maps(materials)
        if textures is not None or uv is not None or maps is not None:
            self._uv = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1)
            self._maps = np.ascontiguousarray(np.zeros((0, 24)) if maps is None else maps, np.float32).reshape(-1, 24)
            if textures is None:
                desc, texels = np.zeros((0, 4)), np.zeros((0, 4))
            elif isinstance(textures, tuple):
                desc, texels = textures
            else:
                desc, texels = pack_textures(textures)
            self._desc = np.ascontiguousarray(desc, np.int32).reshape(-1, 4)
            self._texels = np.ascontiguousarray(texels, np.float32).reshape(-1, 4)
            tx = OraTextures()
            tx.uv = None if self._uv is None else self._uv.ctypes.data
            tx.maps, tx.nmaps = (self._maps.ctypes.data if self._maps.size else None), self._maps.shape[0]
            tx.tex, tx.ntex = (self._desc.ctypes.data if self._desc.size else None), self._desc.shape[0]
            tx.texels = self._texels.ctypes.data if self._texels.size else None
            self._tx = tx
        cfg = OraRenderCfg()
        cfg.width, cfg.height = int(width), int(height)
        cfg.fsaa_x, cfg.fsaa_y = int(samples_per_pixel[0]), int(samples_per_pixel[1])
        cfg.max_bounces, cfg.rr_start, cfg.rr_survive = int(max_bounces), int(min(rr_start, 1 << 30)), float(rr_survive)
        cfg.cull_backfaces = int(ignore_backfaces)
        cfg.seed_lo, cfg.seed_hi = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
        cfg.request, cfg.iteration = 0, 1
        for k in range(13):
            cfg.cam[k] = float(cam_array[k])
        cfg.lens = int(lens)
        cfg.boxing = int(boxing)
        rows = cam_array[3:12] if cam_rows is None else cam_rows
        for k in range(9):
            cfg.cam_rows[k] = float(rows[k])
        if lens_params is not None:
            for k in range(8):
                cfg.lens_p[k] = float(lens_params[k])
        self.cfg = cfg

    def set_albedo(self, i, rgb):
        for c in range(3):
            self.mats[i].albedo[c] = float(np.float32(rgb[c]))

    def set_emission(self, i, rgb):
        for c in range(3):
            self.lights[i].emission[c] = float(np.float32(rgb[c]))

    def render(self, iteration=1, fp64=False, nthreads=0, request=0, rows=None):
        """Image (rows x W x 3); rows = (row0, row1) renders only that band of the frame."""
        self.cfg.iteration = int(iteration)
        self.cfg.request = int(request)
        H, W = self.cfg.height, self.cfg.width
        r0, r1 = (0, H) if rows is None else (int(rows[0]), int(rows[1]))
        out = np.zeros((r1 - r0, W, 3), np.float64 if fp64 else np.float32)
        rc = lib().ora_render(self.tree.h, self.mat_id.ctypes.data, ctypes.addressof(self.mats), self.nmat,
                              ctypes.addressof(self.lights), self.nlights, ctypes.addressof(self.cfg),
                              None if fp64 else out.ctypes.data, out.ctypes.data if fp64 else None, r0, r1,
                              int(nthreads))
        assert rc == 0
        return out

    def render_ref(self, iteration=1, nthreads=0, request=0, rows=None):
        """Image of the reference estimator with the reference's own plugins (ora_render_ref)."""
        self.cfg.iteration = int(iteration)
        self.cfg.request = int(request)
        H, W = self.cfg.height, self.cfg.width
        r0, r1 = (0, H) if rows is None else (int(rows[0]), int(rows[1]))
        out = np.zeros((r1 - r0, W, 3), np.float32)
        nrm = None if self.normals is None else self.normals.ctypes.data
        tx = None if self._tx is None else ctypes.addressof(self._tx)
        rc = lib().ora_render_ref_tex(self.tree.h, self.mat_id.ctypes.data, ctypes.addressof(self.mats), self.nmat,
                                      ctypes.addressof(self.lights), self.nlights, nrm, tx, ctypes.addressof(self.cfg),
                                      out.ctypes.data, r0, r1, int(nthreads))
        assert rc == 0
        return out

    def backward(self, dl_dimage, iteration=1, nthreads=0):
        self.cfg.iteration = int(iteration)
        dl = np.ascontiguousarray(dl_dimage, np.float32).reshape(self.cfg.height, self.cfg.width, 3)
        ga = np.zeros((max(self.nmat, 1), 3), np.float64)
        ge = np.zeros((max(self.nlights, 1), 3), np.float64)
        rc = lib().ora_backward(self.tree.h, self.mat_id.ctypes.data, ctypes.addressof(self.mats), self.nmat,
                                ctypes.addressof(self.lights), self.nlights, ctypes.addressof(self.cfg),
                                dl.ctypes.data, ga.ctypes.data, ge.ctypes.data, int(nthreads))
        assert rc == 0
        return ga[: self.nmat], ge[: self.nlights]


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- primitives
def tri_cast(tri, o, d, clip=False):
    t = ctypes.c_float(0.0)
    p = np.zeros(3, np.float32)
    hit = lib().ora_tri_cast(_f32(tri), _f32(o), _f32(d), int(clip), ctypes.byref(t), p)
    return bool(hit), np.float32(t.value), p


def tri_contains(tri, p):
    return bool(lib().ora_tri_contains(_f32(tri), _f32(p)))


def aabb_cast_preinv(box, o, inv):
    return np.float32(lib().ora_aabb_cast_preinv(_f32(box), _f32(o), _f32(inv)))


def aabb_intersects_tri(box, tri):
    return bool(lib().ora_aabb_intersects_tri(_f32(box), _f32(tri)))


def intersection_bounds(box, tri):
    out = np.zeros(6, np.float32)
    lib().ora_intersection_bounds(_f32(box), _f32(tri), out)
    return out


def intersection_center(box, tri):
    out = np.zeros(3, np.float32)
    lib().ora_intersection_center(_f32(box), _f32(tri), out)
    return out


def clip_tri(box, tri):
    out = np.zeros(27 * 9, np.float32)
    n = lib().ora_clip_tri(_f32(box), _f32(tri), out)
    return out[: 9 * min(n, 27)].reshape(-1, 9), n


def tri_get_masses(tri, p):
    out = np.zeros(3, np.float32)
    lib().ora_tri_get_masses(_f32(tri), _f32(p), out)
    return out


def tri_mass_center(tri, m):
    out = np.zeros(3, np.float32)
    lib().ora_tri_mass_center_w(_f32(tri), _f32(m), out)
    return out


# ---------------------------------------------------------------- octree
class Octree:
    """Oracle octree: Octree<Renderable<BakedTriFace>>::push(all) + build()."""

    def __init__(self, tris, split_threshold=16, max_depth=12):
        self.tris = _f32(tris).reshape(-1, 9)
        self.h = lib().ora_octree_build(self.tris, int(self.tris.shape[0]), int(split_threshold), int(max_depth))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ora_octree_free(self.h)
                self.h = None
        except Exception:            # interpreter shutdown: the module's globals may be gone
            pass

    @property
    def node_count(self):
        return lib().ora_octree_node_count(self.h)

    def export(self):
        n = self.node_count
        tot = lib().ora_octree_list_total(self.h)
        bounds = np.zeros(n * 6, np.float32)
        children = np.zeros(n, np.int32)
        off = np.zeros(n + 1, np.int32)
        lst = np.zeros(max(tot, 1), np.int32)
        lib().ora_octree_export(self.h, bounds, children, off, lst)
        return bounds.reshape(n, 6), children, off, lst[:tot]

    def cast(self, rays, clip=False, exclude=None, counters=False, points=False, nthreads=0):
        rays = _f32(rays).reshape(-1, 6)
        n = rays.shape[0]
        ids = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.int32)
        pts = np.zeros((n, 3), np.float32) if points else None
        nb = np.zeros(n, np.int32) if counters else None
        nr = np.zeros(n, np.int32) if counters else None
        nt = np.zeros(n, np.int32) if counters else None
        lib().ora_octree_cast(self.h, rays, n, int(clip), _ptr(ex), ids, t, _ptr(pts), _ptr(nb), _ptr(nr),
                              _ptr(nt), int(nthreads))
        out = {"id": ids, "t": t}
        if points:
            out["point"] = pts
        if counters:
            out.update(n_box=nb, n_leafref=nr, n_tri=nt)
        return out


def brute_cast(tris, rays, clip=False, exclude=None, nthreads=0):
    tris = _f32(tris).reshape(-1, 9)
    rays = _f32(rays).reshape(-1, 6)
    n = rays.shape[0]
    ids = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    second = np.zeros(n, np.float32)
    ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.int32)
    lib().ora_brute_cast(tris, int(tris.shape[0]), rays, n, int(clip), _ptr(ex), ids, t, _ptr(second), int(nthreads))
    return {"id": ids, "t": t, "second": second}


def frame_bytes(img):
    """Images::translateChannelFloatToByte (Images.cu:7-12) per channel of an (..., 4) float32
    image: v = 255 * c in fp32; 255 when v > 255, 0 when v < 0 (or NaN), else (unsigned char)v."""
    c = np.asarray(img, np.float32)
    v = np.float32(255.0) * c
    out = np.zeros(v.shape, np.uint8)
    ok = (v >= 0) & (v <= 255)
    out[ok] = v[ok].astype(np.uint8)          # truncation toward zero
    out[v > 255] = 255
    return out


def blend(prev, img, k):
    """Iteration k's frame: img * (1/k) + prev * (1 - 1/k) in fp32 (k = 1 sets the frame;
    DumbRenderer.cu:64, BlockBasedFrameBuffer::blendBlockPixelColor)."""
    img = np.asarray(img, np.float32)
    if k == 1 or prev is None:
        return img.copy()
    a = np.float32(1.0) / np.float32(k)
    return (img * a + np.asarray(prev, np.float32) * (np.float32(1.0) - a)).astype(np.float32)
