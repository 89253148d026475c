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
    _fields_ = [("kind", ctypes.c_int32), ("albedo", ctypes.c_float * 3)]


class OraLight(ctypes.Structure):
    _fields_ = [("corner", ctypes.c_float * 3), ("edge_u", ctypes.c_float * 3), ("edge_v", ctypes.c_float * 3),
                ("emission", ctypes.c_float * 3)]


class OraRenderCfg(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("fsaa_x", ctypes.c_int32),
                ("fsaa_y", ctypes.c_int32), ("max_bounces", ctypes.c_int32), ("rr_start", ctypes.c_int32),
                ("rr_survive", ctypes.c_float), ("cull_backfaces", ctypes.c_int32), ("seed_lo", ctypes.c_uint32),
                ("seed_hi", ctypes.c_uint32), ("request", ctypes.c_uint32), ("iteration", ctypes.c_int32),
                ("cam", ctypes.c_float * 13)]


KINDS = {"lambert": 0, "mirror": 1, "metal": 2}


def _bind_render(L):
    vp = ctypes.c_void_p
    L.ora_render.restype = ctypes.c_int
    L.ora_render.argtypes = [vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int]
    L.ora_backward.restype = ctypes.c_int
    L.ora_backward.argtypes = [vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int]
    L.ora_philox.argtypes = [vp, vp, vp]
    L.ora_sincos_2pi.argtypes = [ctypes.c_float, vp, vp]


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
                 max_bounces=2, rr_start=1 << 30, rr_survive=1.0, ignore_backfaces=True, seed=0x5EED):
        self.tree = tree
        self.mat_id = np.ascontiguousarray(mat_id, np.int32)
        self.mats = (OraMaterial * max(len(materials), 1))()
        for i, m in enumerate(materials):
            self.mats[i].kind = KINDS[m["kind"]]
            for c in range(3):
                self.mats[i].albedo[c] = float(np.float32(m["albedo"][c]))
        self.nmat = len(materials)
        self.lights = (OraLight * max(len(lights), 1))()
        for i, l in enumerate(lights):
            for c in range(3):
                self.lights[i].corner[c] = l["corner"][c]
                self.lights[i].edge_u[c] = l["edge_u"][c]
                self.lights[i].edge_v[c] = l["edge_v"][c]
                self.lights[i].emission[c] = l["emission"][c]
        self.nlights = len(lights)
        cfg = OraRenderCfg()
        cfg.width, cfg.height = int(width), int(height)
        cfg.fsaa_x, cfg.fsaa_y = int(samples_per_pixel[0]), int(samples_per_pixel[1])
        cfg.max_bounces, cfg.rr_start, cfg.rr_survive = int(max_bounces), int(min(rr_start, 1 << 30)), float(rr_survive)
        cfg.cull_backfaces = int(ignore_backfaces)
        cfg.seed_lo, cfg.seed_hi = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
        cfg.request, cfg.iteration = 0, 1
        for k in range(13):
            cfg.cam[k] = float(cam_array[k])
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
        if getattr(self, "h", None):
            lib().ora_octree_free(self.h)
            self.h = None

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
