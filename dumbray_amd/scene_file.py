"""Scene files: the reference's ``.dumb`` scene description and the OBJ meshes it names.

``DumbRenderContext`` mirrors the reference's DumbRenderContext (DumbRenderContext.cuh,
DumbRenderContext.cu:295-371): ``buildFromFile`` / ``fromFile`` / ``fromText`` load a
scene through the library's loader (dr_scene_file_*, C++), giving the triangles in push
order, their vertex normals and texture coordinates (BakedTriFace::norm, ::tex) and
material ids, the material, material-map and light tables, the textures of the scene's
PNG files, the camera and the renderer settings (bounces, samples per pixel, boxing,
ignore backfaces, lens).  ``octree()`` builds Octree<Renderable<BakedTriFace>> with the
reference's constants (split threshold 8, maximum depth 24, Octree.cuh:11-13) and
``renderer(width, height)`` a DumbRenderer with the file's settings; the window size
that picks the resolution in the reference is the caller's choice here.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import DumbRayError, check, lib

BOXING_NAMES = {0: "height", 1: "width", 2: "min", 3: "max"}


class SceneFileSettings(ctypes.Structure):
    _fields_ = [("boxing", ctypes.c_int32), ("max_bounces", ctypes.c_int32), ("fsaa_x", ctypes.c_int32),
                ("fsaa_y", ctypes.c_int32), ("ignore_backfaces", ctypes.c_int32), ("lens", ctypes.c_int32),
                ("cam", ctypes.c_float * 13), ("lens_p", ctypes.c_float * 8), ("cam_rows", ctypes.c_float * 9)]


class DumbRenderContext:
    def __init__(self):
        self.tris = np.zeros((0, 9), np.float32)
        self.normals = np.zeros((0, 9), np.float32)
        self.mat_id = np.zeros(0, np.int32)
        self.material_table = np.zeros((0, 12), np.float32)
        self.light_table = np.zeros((0, 16), np.float32)
        self.camera = np.zeros(13, np.float32)
        self.max_bounces, self.samples_per_pixel, self.boxing = 2, (1, 1), "height"
        self.ignore_backfaces, self.lens = True, 0
        self.lens_params = np.zeros(8, np.float32)
        self.cam_rows = np.zeros(9, np.float32)
        self.uv = np.zeros((0, 6), np.float32)
        self.texture_desc = np.zeros((0, 4), np.int32)
        self.texels = np.zeros((0, 4), np.float32)
        self.maps = np.zeros((0, 24), np.float32)

    # ------------------------------------------------------------ loading
    @classmethod
    def fromFile(cls, path: str) -> "DumbRenderContext":
        h = ctypes.c_void_p()
        check(lib().dr_scene_file_load(os.fsencode(path), ctypes.byref(h)), "dr_scene_file_load")
        return cls._take(h)

    @classmethod
    def fromText(cls, text: str, base_dir: str = "") -> "DumbRenderContext":
        h = ctypes.c_void_p()
        check(lib().dr_scene_file_parse(text.encode(), os.fsencode(base_dir), ctypes.byref(h)), "dr_scene_file_parse")
        return cls._take(h)

    buildFromFile = fromFile

    @classmethod
    def _take(cls, h) -> "DumbRenderContext":
        L = lib()
        try:
            st = SceneFileSettings()
            n, nm, nl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            check(L.dr_scene_file_info(h, ctypes.byref(st), ctypes.byref(n), ctypes.byref(nm), ctypes.byref(nl)),
                  "dr_scene_file_info")
            c = cls()
            c.tris = np.zeros((n.value, 9), np.float32)
            c.normals = np.zeros((n.value, 9), np.float32)
            c.mat_id = np.zeros(n.value, np.int32)
            check(L.dr_scene_file_geometry(h, c.tris.ctypes.data, c.normals.ctypes.data, c.mat_id.ctypes.data),
                  "dr_scene_file_geometry")
            c.material_table = np.zeros((nm.value, 12), np.float32)
            c.light_table = np.zeros((nl.value, 16), np.float32)
            check(L.dr_scene_file_tables(h, c.material_table.ctypes.data, c.light_table.ctypes.data),
                  "dr_scene_file_tables")
            c.camera = np.array(st.cam[:], np.float32)
            c.max_bounces = int(st.max_bounces)
            c.samples_per_pixel = (int(st.fsaa_x), int(st.fsaa_y))
            c.boxing = BOXING_NAMES.get(int(st.boxing), "height")
            c.ignore_backfaces = bool(st.ignore_backfaces)
            c.lens = int(st.lens)
            c.lens_params = np.array(st.lens_p[:], np.float32)
            c.cam_rows = np.array(st.cam_rows[:], np.float32)
            c.uv = np.zeros((n.value, 6), np.float32)
            if n.value:
                check(L.dr_scene_file_texcoords(h, c.uv.ctypes.data), "dr_scene_file_texcoords")
            nt, ntx, nm = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int()
            check(L.dr_scene_file_texture_info(h, ctypes.byref(nt), ctypes.byref(ntx), ctypes.byref(nm)),
                  "dr_scene_file_texture_info")
            c.texture_desc = np.zeros((nt.value, 4), np.int32)
            c.texels = np.zeros((ntx.value, 4), np.float32)
            c.maps = np.zeros((nm.value, 24), np.float32)
            check(L.dr_scene_file_textures(h, c.texture_desc.ctypes.data, c.texels.ctypes.data, c.maps.ctypes.data),
                  "dr_scene_file_textures")
            return c
        finally:
            L.dr_scene_file_destroy(h)

    # ------------------------------------------------------------ rendering
    @property
    def triangle_count(self) -> int:
        return int(self.tris.shape[0])

    def octree(self, split_threshold: int = 8, max_depth: int = 24):
        from .octree import Octree

        if self.triangle_count == 0:
            raise DumbRayError("scene file has no geometry")
        t = Octree(split_threshold, max_depth)
        t.push(self.tris)
        t.build()
        return t

    def renderer(self, width: int, height: int, seed: int = 0x5EED, device="cuda", tree=None, **overrides):
        """DumbRenderer over this scene with the file's settings (keyword overrides allowed)."""
        from .render import DumbRenderer

        tree = tree if tree is not None else self.octree()
        kw = dict(samples_per_pixel=self.samples_per_pixel, max_bounces=self.max_bounces,
                  ignore_backfaces=self.ignore_backfaces, lens=self.lens, boxing=self.boxing,
                  lens_params=self.lens_params, cam_rows=self.cam_rows, maps=self.maps,
                  textures=(self.texture_desc, self.texels), uv=self.uv)
        kw.update(overrides)
        return DumbRenderer(tree.upload(), self.mat_id, self.material_table, self.light_table, self.camera, width,
                            height, seed=seed, device=device, normals=self.normals, **kw)
