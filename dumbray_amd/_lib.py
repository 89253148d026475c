"""ctypes binding of libdumbray_hip.so (the C ABI declared in include/dumbray.h).

The HIP library is the only implementation of the hot path: if it is missing
or fails to load, every entry point raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdumbray_hip.so")
_LIB = None


class DumbRayError(RuntimeError):
    pass


def _declare(L):
    c_int, c_uint, c_float, vp, c_char_p = ctypes.c_int, ctypes.c_uint, ctypes.c_float, ctypes.c_void_p, ctypes.c_char_p
    i64 = ctypes.c_int64
    decl = {
        "dr_last_error": (c_char_p, []),
        "dr_version": (c_char_p, []),
        "dr_device_open": (c_int, [c_int]),
        "dr_octree_build": (c_int, [vp, c_int, c_int, c_int, c_int, ctypes.POINTER(vp)]),
        "dr_octree_info": (c_int, [vp, vp, vp, vp, vp]),
        "dr_octree_export": (c_int, [vp, vp, vp, vp, vp]),
        "dr_octree_destroy": (None, [vp]),
        "dr_scene_create": (c_int, [vp, ctypes.POINTER(vp)]),
        "dr_scene_destroy": (None, [vp]),
        "dr_scene_device_bytes": (i64, [vp]),
        "dr_cast": (c_int, [vp, vp, c_int, c_uint, vp, vp, vp, vp, vp]),
    }
    optional = {
        "dr_cast_stats": (c_int, [vp, vp, c_int, c_uint, vp, vp, vp, vp, vp]),
        "dr_scene_set_materials": (c_int, [vp, vp, c_int]),
        "dr_scene_set_normals": (c_int, [vp, vp, c_int]),
        "dr_scene_set_texcoords": (c_int, [vp, vp, c_int]),
        "dr_shading_create": (c_int, [vp, vp, vp, vp, ctypes.POINTER(vp)]),
        "dr_shading_destroy": (None, [vp]),
        "dr_render_blocks": (c_int, [vp, vp, vp, c_int, c_int, vp]),
        "dr_render_stats": (c_int, [vp, vp, vp, c_int, c_int, vp, vp]),
        "dr_backward": (c_int, [vp, vp, vp, vp, vp, vp]),
        "dr_backward_noalloc": (c_int, [vp, vp, vp, vp, vp, vp]),
        "dr_batch_create": (c_int, [vp, c_int, vp, vp]),
        "dr_batch_run": (c_int, [vp, vp]),
        "dr_batch_get_tuning": (c_int, [vp, vp]),
        "dr_batch_destroy": (None, [vp]),
        "dr_scene_file_load": (c_int, [c_char_p, ctypes.POINTER(vp)]),
        "dr_scene_file_parse": (c_int, [c_char_p, c_char_p, ctypes.POINTER(vp)]),
        "dr_scene_file_info": (c_int, [vp, vp, vp, vp, vp]),
        "dr_scene_file_geometry": (c_int, [vp, vp, vp, vp]),
        "dr_scene_file_tables": (c_int, [vp, vp, vp]),
        "dr_scene_file_destroy": (None, [vp]),
        "dr_scene_file_texture_info": (c_int, [vp, vp, vp, vp]),
        "dr_scene_file_textures": (c_int, [vp, vp, vp, vp]),
        "dr_scene_file_texcoords": (c_int, [vp, vp]),
        "dr_frame_to_rgba8": (c_int, [vp, c_int, c_int, vp, vp]),
        "dr_png_write": (c_int, [c_char_p, vp, c_int, c_int]),
        "dr_png_info": (c_int, [c_char_p, vp, vp]),
        "dr_png_read": (c_int, [c_char_p, vp, c_int, c_int]),
    }
    for name, (res, args) in decl.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    for name, (res, args) in optional.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise DumbRayError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _LIB = L
    return _LIB


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().dr_last_error().decode(errors="replace")
        raise DumbRayError(f"{what} failed (status {rc}): {msg}")
