"""Image output (the reference's Images namespace, Images.cuh / Images.cu).

``frame_to_rgba8`` converts a device BlockBasedFrameBuffer<8,16> to a row-major RGBA8
tensor on the device (dr_frame_to_rgba8: 255*c clamped to [0, 255], truncated,
Images.cu:7-19); ``saveBufferPNG`` writes it as an 8-bit RGBA PNG (dr_png_write,
Images.cu:22-36) and returns the reference's error codes."""
from __future__ import annotations

import ctypes
import os

from ._lib import check, lib
from .octree import launch_stream

IMAGES_NO_ERROR = 0
IMAGES_ERROR_FRAME_BUFFER_INVALID = 1
IMAGES_ERROR_ALLOCATION_FAILED = 2
IMAGES_ERROR_EXTERNAL_LIBRARY_FAILED = 3


def frame_to_rgba8(frame, stream=None):
    """(H, W, 4) uint8 device tensor of the frame."""
    import torch

    w, h = frame.getSize()
    out = torch.empty((h, w, 4), dtype=torch.uint8, device=frame.data.device)
    s = launch_stream(stream, frame.data.device, frame.data, out)
    check(lib().dr_frame_to_rgba8(ctypes.c_void_p(frame.data.data_ptr()), int(w), int(h),
                                  ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream)), "dr_frame_to_rgba8")
    return out


def saveBufferPNG(frame, filename: str) -> int:
    w, h = frame.getSize()
    if w <= 0 or h <= 0:
        return IMAGES_ERROR_FRAME_BUFFER_INVALID
    rgba = frame_to_rgba8(frame).cpu().contiguous()
    rc = lib().dr_png_write(os.fsencode(filename), ctypes.c_void_p(rgba.data_ptr()), int(w), int(h))
    return IMAGES_NO_ERROR if rc == 0 else IMAGES_ERROR_EXTERNAL_LIBRARY_FAILED
