"""Host mirror of the reference's geometry container, over the C ABI.

``Octree`` keeps the shape of ``Octree<Renderable<BakedTriFace>>``
(Octree.cuh:50-228): ``reset``/``push``/``build``/``getNodeCount``/``getData``
and ``cast``.  ``push`` collects triangles in order (their index is the hit id,
Octree.impl.cuh:97-111); ``build`` runs the native parallel builder
(dr_octree_build); ``upload`` makes the device-resident scene (replacing the
deep copy of TypeTools<Octree>::prepareForCpyLoad, Octree.impl.cuh:763-796);
``cast`` is batch ``Octree::cast`` (Octree.impl.cuh:174-212) on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import DumbRayError, check, lib


def launch_stream(stream, device, *tensors):
    """The stream a library call launches on: torch's current stream, or `stream`, which then
    first waits for the work queued on the current stream (where the call's input and output
    tensors were made), and keeps those tensors' memory from reuse until its work is done."""
    import torch

    cur = torch.cuda.current_stream(device)
    if stream is None or stream == cur:
        return cur
    stream.wait_stream(cur)
    for t in tensors:
        if t is not None:
            t.record_stream(stream)
    return stream

OCTREE_POLYCOUNT_TO_SPLIT_NODE = 8   # Octree.cuh:11
OCTREE_MAX_DEPTH = 24                # Octree.cuh:13
DR_CAST_CLIP_BACKFACES = 1


def _vp(x):
    return ctypes.c_void_p(x)


class Octree:
    def __init__(self, split_threshold: int = OCTREE_POLYCOUNT_TO_SPLIT_NODE, max_depth: int = OCTREE_MAX_DEPTH):
        self.split_threshold = int(split_threshold)
        self.max_depth = int(max_depth)
        self._parts: list[np.ndarray] = []
        self._data = np.zeros((0, 9), np.float32)
        self._h = None

    # ------------------------------------------------------------ push & build
    def reset(self):
        self._free()
        self._parts = []
        self._data = np.zeros((0, 9), np.float32)

    def push(self, tris):
        t = np.ascontiguousarray(np.asarray(tris, dtype=np.float32)).reshape(-1, 9)
        self._parts.append(t)
        self._free()

    def build(self, nthreads: int = 0):
        if self._parts:
            self._data = np.ascontiguousarray(np.concatenate([self._data] + self._parts, axis=0))
            self._parts = []
        self._free()
        h = ctypes.c_void_p()
        rc = lib().dr_octree_build(self._data.ctypes.data_as(ctypes.c_void_p), int(self._data.shape[0]),
                                   self.split_threshold, self.max_depth, int(nthreads), ctypes.byref(h))
        check(rc, "dr_octree_build")
        self._h = h
        return self

    def getData(self) -> np.ndarray:
        return self._data

    def _info(self):
        self._require()
        vals = [ctypes.c_int() for _ in range(4)]
        check(lib().dr_octree_info(self._h, *[ctypes.byref(v) for v in vals]), "dr_octree_info")
        return [v.value for v in vals]

    def getNodeCount(self) -> int:
        return self._info()[0]

    def device_layout_sizes(self):
        n, tot, groups, recs = self._info()
        return {"nodes": n, "list_total": tot, "groups": groups, "records": recs}

    def export(self):
        """(bounds[n,6], children[n], list_off[n+1], list[total]) in breadth-first node order."""
        n, tot, _, _ = self._info()
        bounds = np.zeros((n, 6), np.float32)
        children = np.zeros(n, np.int32)
        off = np.zeros(n + 1, np.int32)
        lst = np.zeros(max(tot, 1), np.int32)
        check(lib().dr_octree_export(self._h, bounds.ctypes.data_as(ctypes.c_void_p),
                                     children.ctypes.data_as(ctypes.c_void_p), off.ctypes.data_as(ctypes.c_void_p),
                                     lst.ctypes.data_as(ctypes.c_void_p)), "dr_octree_export")
        return bounds, children, off, lst[:tot]

    def upload(self) -> "DeviceScene":
        self._require()
        return DeviceScene(self)

    # ------------------------------------------------------------------ misc
    def _require(self):
        if self._h is None:
            raise DumbRayError("Octree.build() has not been called")

    def _free(self):
        if self._h is not None:
            lib().dr_octree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


class DeviceScene:
    """Device-resident octree (+ shading tables once attached)."""

    def __init__(self, octree: Octree):
        h = ctypes.c_void_p()
        check(lib().dr_scene_create(octree._h, ctypes.byref(h)), "dr_scene_create")
        self._h = h
        self.ntris = int(octree.getData().shape[0])
        self._shadings = {}

    def shading(self, mat_id, normals=None, uv=None) -> "Shading":
        """Shading tables of one renderer of this scene (dr_shading_create): material ids (ntris,),
        vertex normals (ntris, 9) or None for flat normals, texture coordinates (ntris, 6) or None
        for zeros.  Renderers with equal tables share one handle."""
        import hashlib
        import weakref

        mid = np.ascontiguousarray(np.asarray(mat_id, np.int32).reshape(-1))
        nrm = None if normals is None else np.ascontiguousarray(np.asarray(normals, np.float32).reshape(-1, 9))
        tex = None if uv is None else np.ascontiguousarray(np.asarray(uv, np.float32).reshape(-1, 6))
        if mid.shape[0] != self.ntris or (nrm is not None and nrm.shape[0] != self.ntris) or \
                (tex is not None and tex.shape[0] != self.ntris):
            raise DumbRayError("shading: one material id (and normal / texture coordinate row) per triangle required")
        key = hashlib.sha1(b"".join([mid.tobytes(), b"n" if nrm is None else nrm.tobytes(),
                                     b"u" if tex is None else tex.tobytes()])).hexdigest()
        ref = self._shadings.get(key)
        sh = ref() if ref is not None else None
        if sh is None:
            sh = Shading(self, mid, nrm, tex)
            self._shadings[key] = weakref.ref(sh)
        return sh

    @property
    def handle(self):
        return self._h

    def device_bytes(self) -> int:
        return int(lib().dr_scene_device_bytes(self._h))

    def set_normals(self, normals=None):
        """Per-vertex normals of every triangle (BakedTriFace::norm): (ntris, 9) float32 or
        None for flat normals, the unit Triangle::normal (c - a) x (b - a)."""
        if normals is None:
            check(lib().dr_scene_set_normals(self._h, None, self.ntris), "dr_scene_set_normals")
            return
        n = np.ascontiguousarray(np.asarray(normals, np.float32).reshape(-1, 9))
        if n.shape[0] != self.ntris:
            raise DumbRayError("set_normals: one row of 9 floats per triangle required")
        check(lib().dr_scene_set_normals(self._h, n.ctypes.data_as(ctypes.c_void_p), self.ntris),
              "dr_scene_set_normals")

    def set_texcoords(self, uv=None):
        """Texture coordinates of every triangle (BakedTriFace::tex): (ntris, 6) float32 (u, v of
        a, b, c) or None for zeros."""
        if uv is None:
            check(lib().dr_scene_set_texcoords(self._h, None, self.ntris), "dr_scene_set_texcoords")
            return
        t = np.ascontiguousarray(np.asarray(uv, np.float32).reshape(-1, 6))
        if t.shape[0] != self.ntris:
            raise DumbRayError("set_texcoords: one row of 6 floats per triangle required")
        check(lib().dr_scene_set_texcoords(self._h, t.ctypes.data_as(ctypes.c_void_p), self.ntris),
              "dr_scene_set_texcoords")

    def cast(self, rays, clip_backfaces: bool = True, exclude=None, points: bool = False, stream=None):
        """Batch closest hit.  rays: torch float32 (n, 6) on the GPU.  Returns dict of
        torch tensors id (int32, -1 = miss), t (float32, FLT_MAX on miss) [, point]."""
        import torch

        if not (isinstance(rays, torch.Tensor) and rays.is_cuda and rays.dtype == torch.float32):
            raise DumbRayError("cast: rays must be a float32 GPU tensor of shape (n, 6)")
        rays = rays.contiguous().view(-1, 6)
        n = rays.shape[0]
        ids = torch.empty(n, dtype=torch.int32, device=rays.device)
        t = torch.empty(n, dtype=torch.float32, device=rays.device)
        pts = torch.empty((n, 3), dtype=torch.float32, device=rays.device) if points else None
        ex = None
        if exclude is not None:
            ex = exclude.contiguous().to(device=rays.device, dtype=torch.int32).view(-1)
            if ex.numel() != n:
                raise DumbRayError("cast: exclude must hold one triangle id per ray")
        s = launch_stream(stream, rays.device, rays, ex, ids, t, pts)
        rc = lib().dr_cast(self._h, _vp(rays.data_ptr()), int(n), DR_CAST_CLIP_BACKFACES if clip_backfaces else 0,
                           _vp(ex.data_ptr()) if ex is not None else None, _vp(ids.data_ptr()), _vp(t.data_ptr()),
                           _vp(pts.data_ptr()) if pts is not None else None, _vp(s.cuda_stream))
        check(rc, "dr_cast")
        out = {"id": ids, "t": t}
        if points:
            out["point"] = pts
        return out

    def cast_stats(self, rays, clip_backfaces: bool = True, stream=None):
        """dr_cast_stats: ids, t and per-ray counters (groups expanded, leaves, plane tests,
        inside tests, live children, slab tests) of the traversal the kernel ran."""
        import torch

        rays = rays.contiguous().view(-1, 6)
        n = rays.shape[0]
        ids = torch.empty(n, dtype=torch.int32, device=rays.device)
        t = torch.empty(n, dtype=torch.float32, device=rays.device)
        cnt = torch.empty((n, 6), dtype=torch.int32, device=rays.device)
        s = launch_stream(stream, rays.device, rays, ids, t, cnt)
        rc = lib().dr_cast_stats(self._h, _vp(rays.data_ptr()), int(n),
                                 DR_CAST_CLIP_BACKFACES if clip_backfaces else 0, None, _vp(ids.data_ptr()),
                                 _vp(t.data_ptr()), _vp(cnt.data_ptr()), _vp(s.cuda_stream))
        check(rc, "dr_cast_stats")
        return {"id": ids, "t": t, "counts": cnt}

    def __del__(self):
        try:
            if self._h is not None:
                lib().dr_scene_destroy(self._h)
                self._h = None
        except Exception:
            pass


class Shading:
    """dr_shading handle: one renderer's material ids, vertex normals and texture coordinates of a
    DeviceScene, in device memory of its own (dr_render_params.shading)."""

    def __init__(self, scene: DeviceScene, mat_id, normals=None, uv=None):
        h = ctypes.c_void_p()
        ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        check(lib().dr_shading_create(scene.handle, ptr(mat_id), ptr(normals), ptr(uv), ctypes.byref(h)),
              "dr_shading_create")
        self._h = h
        self.scene = scene
        self.max_mat = int(mat_id.max()) if mat_id.size else -1

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h is not None:
                lib().dr_shading_destroy(self._h)
                self._h = None
        except Exception:
            pass
