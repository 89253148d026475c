"""dumbray_amd: MI355X-native implementation of DumbRay's path-tracing hot path.

Octree closest-hit traversal and the path-tracing loop run as hand-written HIP
kernels for gfx950 in libdumbray_hip.so, reached through the C ABI in
include/dumbray.h.  This package is the host-side mirror of the reference's
plugin/operator interface over that ABI.
"""
from ._lib import DumbRayError, lib  # noqa: F401
from .octree import OCTREE_MAX_DEPTH, OCTREE_POLYCOUNT_TO_SPLIT_NODE, DeviceScene, Octree  # noqa: F401
from . import plugins  # noqa: F401
from .render import BlockBasedFrameBuffer, DumbRenderer, RenderBatch  # noqa: F401
from .scene_file import DumbRenderContext  # noqa: F401
from .render_process import BufferedRenderProcess  # noqa: F401
from . import images  # noqa: F401

__all__ = ["DumbRayError", "lib", "Octree", "DeviceScene", "OCTREE_MAX_DEPTH", "OCTREE_POLYCOUNT_TO_SPLIT_NODE",
           "BlockBasedFrameBuffer", "DumbRenderer", "RenderBatch", "DumbRenderContext", "BufferedRenderProcess", "images", "plugins"]
