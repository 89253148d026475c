"""The C-ABI library loads and exports every function declared in include/dumbray.h
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import dumbray_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "dumbray.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(dr_[a-z0-9_]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("dr_cast", "dr_octree_build", "dr_scene_create", "dr_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(dumbray_amd._lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_error_strings():
    L = dumbray_amd.lib()
    assert b"gfx950" in L.dr_version()
    assert isinstance(L.dr_last_error(), bytes)


def test_error_path_without_a_device():
    """A refused call returns DR_ERR_ARG and names the problem in dr_last_error() before it
    touches the device, so a caller can fall back the way the reference does after a failed
    renderBlocksGPU (DumbRenderer.cu:109-111)."""
    import dumbray_amd.render as R
    L = dumbray_amd.lib()
    p = R.RenderParams()
    assert L.dr_render_blocks(None, ctypes.byref(p), None, 0, 1, None) == 1
    assert b"null scene" in L.dr_last_error()
    assert L.dr_backward(None, ctypes.byref(p), None, None, None, None) == 1
    assert b"null scene" in L.dr_last_error()
    assert L.dr_batch_create(None, 0, None, ctypes.byref(ctypes.c_void_p())) == 1
    assert b"dr_batch_create" in L.dr_last_error()
    assert L.dr_shading_create(None, None, None, None, ctypes.byref(ctypes.c_void_p())) == 1
    assert b"dr_shading_create" in L.dr_last_error()
    assert L.dr_cast(None, None, 0, 0, None, None, None, None, None) == 1
    assert b"dr_cast" in L.dr_last_error()
