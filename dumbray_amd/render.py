"""Host mirror of the reference's render operator over the C ABI.

``BlockBasedFrameBuffer`` keeps the reference's ``BlockBasedFrameBuffer<8,16>``
layout (BlockBasedFrameBuffer.cuh:34-91, .impl.cuh:108-156): RGBA fp32, block-
major, 128 pixels per block, in device memory.  ``DumbRenderer`` keeps the
settings of ``DumbRenderer`` (DumbRenderer.cuh:20-31: max bounces, samples per
pixel X/Y, ignore backfaces; HEIGHT boxing) and its ``renderBlocksGPU``
(DumbRenderer.cu:87-113), plus ``backward`` (the adjoint, dr_backward).
Materials and lights live in device tables owned by the renderer, so
``albedo`` and ``emission`` are torch views that an optimiser can update in
place between renders.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import plugins as PL
from ._lib import DumbRayError, check, lib
from .octree import DeviceScene, launch_stream

DR_LAMBERT, DR_MIRROR, DR_METAL = 0, 1, 2
KINDS = PL.MATERIAL_KINDS
LENSES = {"stochastic": PL.STOCHASTIC_LENS, "perspective": PL.PERSPECTIVE_LENS,
          "spherical_segment": PL.SPHERICAL_SEGMENT_LENS}
BLOCK_W, BLOCK_H = 8, 16
BLOCK_SIZE = BLOCK_W * BLOCK_H
DUMB_RENDERER_BOUNCE_LIMIT = 64      # DumbRenderer.cuh:7


class RenderParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("fsaa_x", ctypes.c_int32),
                ("fsaa_y", ctypes.c_int32), ("max_bounces", ctypes.c_int32), ("rr_start", ctypes.c_int32),
                ("rr_survive", ctypes.c_float), ("cull_backfaces", ctypes.c_int32), ("seed_lo", ctypes.c_uint32),
                ("seed_hi", ctypes.c_uint32), ("request", ctypes.c_uint32), ("iteration", ctypes.c_int32),
                ("cam", ctypes.c_float * 13), ("materials", ctypes.c_void_p), ("lights", ctypes.c_void_p),
                ("nmat", ctypes.c_int32), ("nlights", ctypes.c_int32), ("lens", ctypes.c_int32),
                ("plugins", ctypes.c_int32), ("boxing", ctypes.c_int32), ("lens_p", ctypes.c_float * 8),
                ("cam_rows", ctypes.c_float * 9), ("maps", ctypes.c_void_p), ("textures", ctypes.c_void_p),
                ("texels", ctypes.c_void_p), ("nmaps", ctypes.c_int32), ("ntextures", ctypes.c_int32),
                ("shading", ctypes.c_void_p)]


BOXING = {"height": 0, "width": 1, "min": 2, "max": 3}   # DumbRenderer::BoxingMode (DumbRenderer.cuh:12-18)


def _vp(x):
    return ctypes.c_void_p(x)


def block_count(width: int, height: int) -> int:
    return ((width + BLOCK_W - 1) // BLOCK_W) * ((height + BLOCK_H - 1) // BLOCK_H)


def block_to_image_index(width: int, height: int) -> np.ndarray:
    """For every row-major pixel (y, x), its row in the block-major buffer
    (BlockBasedFrameBuffer::pixelBlockLocation, BlockBasedFrameBuffer.impl.cuh:108-123)."""
    nbw = (width + BLOCK_W - 1) // BLOCK_W
    y, x = np.mgrid[0:height, 0:width]
    block = (y // BLOCK_H) * nbw + (x // BLOCK_W)
    pixel = (y % BLOCK_H) * BLOCK_W + (x % BLOCK_W)
    return (block * BLOCK_SIZE + pixel).reshape(-1)


class BlockBasedFrameBuffer:
    """Device BlockBasedFrameBuffer<8,16>: data is a float32 tensor (blocks*128, 4)."""

    def __init__(self, width: int, height: int, device="cuda"):
        self.device = device
        self.setResolution(width, height)

    def setResolution(self, width: int, height: int) -> bool:
        """FrameBuffer::setResolution: a new, zeroed frame of the given size."""
        import torch

        self.width, self.height = int(width), int(height)
        self.data = torch.zeros((block_count(self.width, self.height) * BLOCK_SIZE, 4), dtype=torch.float32,
                                device=self.device)
        self._idx = torch.from_numpy(block_to_image_index(self.width, self.height)).to(self.data.device)
        return True

    def getBlockSize(self) -> int:
        return BLOCK_SIZE

    def getBlockCount(self) -> int:
        return block_count(self.width, self.height)

    def getSize(self):
        return self.width, self.height

    def to_image(self):
        """(H, W, 4) row-major view of the frame."""
        return self.data[self._idx].view(self.height, self.width, 4)

    def from_image(self, img):
        """Block-major (blocks*128, 4) tensor from an (H, W, C) image (C = 3 or 4)."""
        import torch

        img = img.reshape(self.height * self.width, -1)
        out = torch.zeros_like(self.data)
        out[self._idx, : img.shape[1]] = img.to(out.dtype)
        return out


def _tables(materials, lights):
    """(material table (n,12), light table (n,16), nmat, nlights, plugin set) from lists of
    plugin dicts or from tables already in dr_material / dr_light layout (scene files)."""
    if isinstance(materials, np.ndarray):
        mt = np.ascontiguousarray(materials, np.float32).reshape(-1, 12)
        nmat = mt.shape[0]
        mkinds = mt[:, 0].copy().view(np.int32)
        mt = mt if nmat else np.zeros((1, 12), np.float32)
    else:
        mt, nmat = PL.material_table(materials), len(materials)
        mkinds = np.array([PL.MATERIAL_KINDS[m["kind"]] for m in materials], np.int32)
    if isinstance(lights, np.ndarray):
        lt = np.ascontiguousarray(lights, np.float32).reshape(-1, 16)
        nlights = lt.shape[0]
        lkinds = lt[:, 12].copy().view(np.int32)
        lt = lt if nlights else np.zeros((1, 16), np.float32)
    else:
        lt, nlights = PL.light_table(lights), len(lights)
        lkinds = np.array([PL.LIGHT_KINDS[l.get("kind", "quad")] for l in lights], np.int32)
    reference = bool((mkinds >= PL.MATERIAL_KINDS["dumb_based"]).any() or (lkinds != 0).any())
    return mt, lt, nmat, nlights, (PL.PLUGINS_REFERENCE if reference else PL.PLUGINS_BENCHMARK)


class DumbRenderer:
    """``materials`` / ``lights``: lists of plugin dicts (dumbray_amd.plugins) or tables in
    dr_material (n,12) / dr_light (n,16) float32 layout; ``camera``: a Camera or 13 floats
    (dr_render_params.cam); ``boxing``: "height", "width", "min" or "max"."""

    def __init__(self, scene: DeviceScene, mat_id, materials, lights, camera, width: int, height: int,
                 samples_per_pixel=(1, 1), max_bounces: int = 2, rr_start: int = 1 << 30, rr_survive: float = 1.0,
                 ignore_backfaces: bool = True, seed: int = 0x5EED, device="cuda", lens: str = "stochastic",
                 normals=None, boxing="height", lens_params=None, cam_rows=None, textures=None, uv=None,
                 maps=None):
        import torch

        self.scene = scene
        self.width, self.height = int(width), int(height)
        self.fsaa_x, self.fsaa_y = (max(1, int(samples_per_pixel[0])), max(1, int(samples_per_pixel[1])))
        self.max_bounces = min(int(max_bounces), DUMB_RENDERER_BOUNCE_LIMIT)
        self.rr_start, self.rr_survive = int(min(rr_start, 1 << 30)), float(rr_survive)
        self.ignore_backfaces = bool(ignore_backfaces)
        self.seed = int(seed)
        self.camera = camera
        # this renderer's material ids, vertex normals and texture coordinates (the scene itself is
        # shared with other renderers and stays untouched)
        self.shading = scene.shading(mat_id, normals, uv)
        self.lens = LENSES[lens] if isinstance(lens, str) else int(lens)
        self.boxing = BOXING[boxing] if isinstance(boxing, str) else int(boxing)
        # SphericalSegmentLense parameters (plugins.spherical_segment_lens) and the camera's point-map
        # rows (default: right, up, forward of an unscaled camera)
        self.lens_params = np.zeros(8, np.float32) if lens_params is None else np.asarray(lens_params, np.float32)
        self.cam_rows = None if cam_rows is None else np.asarray(cam_rows, np.float32)
        mt, lt, self.nmat, self.nlights, self.plugins = _tables(materials, lights)
        self.material_table = torch.from_numpy(mt.copy()).to(device)
        self.light_table = torch.from_numpy(lt.copy()).to(device)
        # textures: a list of plugins.texture dicts or (descriptors, texels) tables; maps from the
        # material dicts, or given as a dr_material_maps table with material tables
        if maps is None:
            maps = PL.maps_table(materials) if not isinstance(materials, np.ndarray) else np.zeros((0, 24), np.float32)
        maps = np.ascontiguousarray(maps, np.float32).reshape(-1, 24)
        if textures is None:
            desc, texels = np.zeros((0, 4), np.int32), np.zeros((0, 4), np.float32)
        elif isinstance(textures, tuple):
            desc, texels = (np.ascontiguousarray(textures[0], np.int32).reshape(-1, 4),
                            np.ascontiguousarray(textures[1], np.float32).reshape(-1, 4))
        else:
            desc, texels = PL.texture_tables(textures)
        self.nmaps, self.ntextures = int(maps.shape[0]), int(desc.shape[0])
        self.maps_table = torch.from_numpy(maps.copy() if maps.size else np.zeros((1, 24), np.float32)).to(device)
        self.texture_table = torch.from_numpy(desc.copy() if desc.size else np.zeros((1, 4), np.int32)).to(device)
        self.texels = torch.from_numpy(texels.copy() if texels.size else np.zeros((1, 4), np.float32)).to(device)

    # ------------------------------------------------ BufferedRenderer / Renderer iteration interface
    # (Renderer.cu:136-156, BufferedRenderer.cu:7): iterate() renders iteration k + 1 into the
    # frame buffer, blending 1/(k + 1); the resolution follows the frame buffer.
    def setFrameBuffer(self, frame):
        if frame is not getattr(self, "_frame", None):
            self._frame = frame
            self.resetIterations()

    def getFrameBuffer(self):
        return getattr(self, "_frame", None)

    def iteration(self) -> int:
        return getattr(self, "_iteration", 0)

    def resetIterations(self):
        self._iteration = 0

    def iterate(self) -> bool:
        import torch

        frame = self.getFrameBuffer()
        if frame is None:
            return False
        self.width, self.height = frame.getSize()
        k = self.iteration() + 1
        try:
            self.render(frame, k)
            torch.cuda.current_stream(frame.data.device).synchronize()
        except DumbRayError:
            return False
        self._iteration = k
        return True

    # ------------------------------------------------ progress checkpoints (no counterpart in the
    # reference, whose progress is only the in-memory iteration counter and frame, Renderer.cu:137-141)
    def _progress_key(self):
        """The settings that decide which samples iteration k draws and how they blend."""
        rr = np.frombuffer(np.float32(self.rr_survive).tobytes(), np.int32)[0]
        return np.array([self.width, self.height, self.fsaa_x, self.fsaa_y, self.max_bounces, self.rr_start, rr,
                         int(self.ignore_backfaces), self.seed & 0xFFFFFFFF, (self.seed >> 32) & 0xFFFFFFFF, self.lens,
                         self.plugins, self.boxing], np.int64)

    def saveProgress(self, path):
        """Writes the iteration count, the frame buffer (block-major RGBA fp32) and the settings
        that decide the samples to an .npz file.  The scene, materials, lights and camera are
        the caller's to restore.  Iterations are keyed by their number (Philox counters), so a
        renderer that loads the file and iterates n more times ends on the frame of an
        uninterrupted run."""
        frame = self.getFrameBuffer()
        if frame is None:
            raise DumbRayError("saveProgress: no frame buffer")
        np.savez(path, iteration=np.int64(self.iteration()), frame=frame.data.cpu().numpy(), key=self._progress_key())

    def loadProgress(self, path, device="cuda"):
        """Restores a saveProgress file into a new frame buffer of this renderer (which becomes
        its frame buffer) and returns it.  Raises when the file's settings differ."""
        import torch

        with np.load(path, allow_pickle=False) as z:
            it, data, key = int(z["iteration"]), z["frame"], z["key"]
        # the whole key, resolution included, is compared before anything of this renderer changes
        if not np.array_equal(key, self._progress_key()):
            raise DumbRayError("loadProgress: the file was written with other render settings")
        frame = BlockBasedFrameBuffer(self.width, self.height, device)
        if tuple(frame.data.shape) != tuple(data.shape):
            raise DumbRayError("loadProgress: frame size differs from the settings")
        frame.data.copy_(torch.from_numpy(data))
        self.setFrameBuffer(frame)
        self._iteration = it
        return frame

    # parameters an optimiser may update in place
    @property
    def albedo(self):
        return self.material_table[: self.nmat, 1:4]

    @property
    def emission(self):
        return self.light_table[: self.nlights, 9:12]

    def params(self, iteration: int = 1, request: int = 0) -> RenderParams:
        """dr_render_params of this renderer (request = Philox counter word 3)."""
        p = RenderParams()
        p.width, p.height, p.fsaa_x, p.fsaa_y = self.width, self.height, self.fsaa_x, self.fsaa_y
        p.max_bounces, p.rr_start, p.rr_survive = self.max_bounces, self.rr_start, self.rr_survive
        p.cull_backfaces = int(self.ignore_backfaces)
        p.seed_lo, p.seed_hi = self.seed & 0xFFFFFFFF, (self.seed >> 32) & 0xFFFFFFFF
        p.request, p.iteration = int(request), int(iteration)
        cam = self.camera.as_array() if hasattr(self.camera, "as_array") else np.asarray(self.camera, np.float32)
        for k in range(13):
            p.cam[k] = float(cam[k])
        rows = self.cam_rows if self.cam_rows is not None else cam[3:12]
        for k in range(9):
            p.cam_rows[k] = float(rows[k])
        for k in range(8):
            p.lens_p[k] = float(self.lens_params[k])
        p.materials = self.material_table.data_ptr()
        p.lights = self.light_table.data_ptr()
        p.nmat, p.nlights = self.nmat, self.nlights
        p.lens, p.plugins, p.boxing = self.lens, self.plugins, self.boxing
        p.maps, p.nmaps = self.maps_table.data_ptr(), self.nmaps
        p.textures, p.texels, p.ntextures = self.texture_table.data_ptr(), self.texels.data_ptr(), self.ntextures
        p.shading = self.shading.handle.value
        return p

    def renderBlocksGPU(self, frame: BlockBasedFrameBuffer, start_block: int, end_block: int, iteration: int = 1,
                        stream=None):
        import torch

        if frame.getSize() != (self.width, self.height):
            raise DumbRayError("renderBlocksGPU: frame size differs from the renderer's")
        p = self.params(iteration)
        s = launch_stream(stream, frame.data.device, frame.data)
        check(lib().dr_render_blocks(self.scene.handle, ctypes.byref(p), _vp(frame.data.data_ptr()), int(start_block),
                                     int(end_block), _vp(s.cuda_stream)), "dr_render_blocks")

    RENDER_COUNTERS = ("segments", "photons", "group_expansions", "leaves", "plane_tests", "inside_tests", "paths",
                       "live_children", "slab_tests")

    def render_stats(self, frame: BlockBasedFrameBuffer | None = None, iteration: int = 1, stream=None):
        """dr_render_stats: renders like render() and returns (frame, counters of the work the
        kernel did: casts, photon casts, group expansions, leaves, plane and inside tests, paths)."""
        import torch

        if frame is None:
            frame = BlockBasedFrameBuffer(self.width, self.height, self.material_table.device)
        counts = torch.zeros(len(self.RENDER_COUNTERS), dtype=torch.int64, device=frame.data.device)
        p = self.params(iteration)
        s = launch_stream(stream, frame.data.device, frame.data, counts)
        check(lib().dr_render_stats(self.scene.handle, ctypes.byref(p), _vp(frame.data.data_ptr()), 0,
                                    frame.getBlockCount(), _vp(counts.data_ptr()), _vp(s.cuda_stream)),
              "dr_render_stats")
        s.synchronize()
        return frame, dict(zip(self.RENDER_COUNTERS, (int(v) for v in counts.cpu())))

    def render(self, frame: BlockBasedFrameBuffer | None = None, iteration: int = 1, stream=None):
        if frame is None:
            frame = BlockBasedFrameBuffer(self.width, self.height, self.material_table.device)
        self.renderBlocksGPU(frame, 0, frame.getBlockCount(), iteration, stream)
        return frame

    def backward(self, dl_dframe, iteration: int = 1, stream=None):
        """Renderer::backward: (dL/dalbedo (nmat,3), dL/demission (nlights,3)) as float64 tensors for
        the loss gradient dl_dframe (block-major (blocks*128, 4), alpha ignored)."""
        import torch

        dl = dl_dframe.contiguous().to(torch.float32)
        if dl.shape != (block_count(self.width, self.height) * BLOCK_SIZE, 4):
            raise DumbRayError("backward: dl_dframe must have the frame's block layout")
        ga = torch.zeros((self.nmat, 3), dtype=torch.float64, device=dl.device)
        ge = torch.zeros((max(self.nlights, 1), 3), dtype=torch.float64, device=dl.device)
        p = self.params(iteration)
        s = launch_stream(stream, dl.device, dl, ga, ge)
        check(lib().dr_backward(self.scene.handle, ctypes.byref(p), _vp(dl.data_ptr()), _vp(ga.data_ptr()),
                                _vp(ge.data_ptr()), _vp(s.cuda_stream)), "dr_backward")
        return ga, ge[: self.nlights]


class Request(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("params", RenderParams), ("frame", ctypes.c_void_p)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("tile_blocks", ctypes.c_int32), ("groups_per_cu", ctypes.c_int32), ("refill_min", ctypes.c_int32)]


class RenderBatch:
    """Many render requests in one persistent launch captured in a hipGraph (dr_batch_*).

    requests: list of (DumbRenderer, BlockBasedFrameBuffer, iteration, request_id).  The
    renderers' parameters are read when the batch is created.  tuning: None runs the
    autotuner, else a dict with tile_blocks, groups_per_cu, refill_min."""

    def __init__(self, requests, tuning=None):
        self._keep = list(requests)
        arr = (Request * len(requests))()
        for i, (r, frame, iteration, req_id) in enumerate(requests):
            if frame.getSize() != (r.width, r.height):
                raise DumbRayError("RenderBatch: frame size differs from its renderer's")
            arr[i].scene = r.scene.handle.value
            arr[i].params = r.params(iteration, req_id)
            arr[i].frame = frame.data.data_ptr()
        t = None
        if tuning is not None:
            t = BatchTuning(int(tuning["tile_blocks"]), int(tuning["groups_per_cu"]), int(tuning["refill_min"]))
        h = ctypes.c_void_p()
        check(lib().dr_batch_create(arr, len(requests), ctypes.byref(t) if t is not None else None, ctypes.byref(h)),
              "dr_batch_create")
        self._h = h

    @property
    def tuning(self):
        t = BatchTuning()
        check(lib().dr_batch_get_tuning(self._h, ctypes.byref(t)), "dr_batch_get_tuning")
        return {"tile_blocks": t.tile_blocks, "groups_per_cu": t.groups_per_cu, "refill_min": t.refill_min}

    def run(self, stream=None):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        check(lib().dr_batch_run(self._h, _vp(s.cuda_stream)), "dr_batch_run")

    def __del__(self):
        try:
            if self._h is not None:
                lib().dr_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass
