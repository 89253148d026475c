"""Seeded synthetic scenes and ray batches for the benchmark workloads.

Triangles are returned as float32 arrays of shape (n, 9) = (a, b, c) in push
order; the triangle's index in that array is its hit ID (the index into
``Octree::getData()``, Octree.impl.cuh:97-111).

Winding: the reference treats a face as front-facing when the ray origin lies
on the -n side of n = (b - a) x (c - a) (Triangle.impl.h:334-340, backface clip
if (a - o).n < -eps).  Every generator here winds faces so that n points *into*
closed objects and *down* for ground/terrain, so rays arriving from outside or
from above see front faces.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ----------------------------------------------------------------------------
# noise
# ----------------------------------------------------------------------------


def _value_noise_2d(x: np.ndarray, z: np.ndarray, freq: int, rng: np.random.Generator) -> np.ndarray:
    """Lattice value noise on a (freq+1)^2 grid of U(-1,1) values, smoothstep-interpolated."""
    lat = rng.uniform(-1.0, 1.0, size=(freq + 1, freq + 1))
    fx = x * freq
    fz = z * freq
    ix = np.clip(np.floor(fx).astype(np.int64), 0, freq - 1)
    iz = np.clip(np.floor(fz).astype(np.int64), 0, freq - 1)
    tx = fx - ix
    tz = fz - iz
    sx = tx * tx * (3.0 - 2.0 * tx)
    sz = tz * tz * (3.0 - 2.0 * tz)
    v00 = lat[ix, iz]
    v10 = lat[ix + 1, iz]
    v01 = lat[ix, iz + 1]
    v11 = lat[ix + 1, iz + 1]
    a = v00 + (v10 - v00) * sx
    b = v01 + (v11 - v01) * sx
    return a + (b - a) * sz


def fbm_2d(x, z, octaves=6, lacunarity=2.0, gain=0.5, base_freq=4, seed=0):
    rng = np.random.default_rng(seed)
    out = np.zeros(np.broadcast(x, z).shape, dtype=np.float64)
    amp = 1.0
    freq = float(base_freq)
    norm = 0.0
    for _ in range(octaves):
        out += amp * _value_noise_2d(x, z, int(round(freq)), rng)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return out / norm


def _value_noise_3d(p: np.ndarray, freq: int, rng: np.random.Generator) -> np.ndarray:
    lat = rng.uniform(-1.0, 1.0, size=(freq + 1, freq + 1, freq + 1))
    f = np.clip(p, 0.0, 1.0) * freq
    i = np.clip(np.floor(f).astype(np.int64), 0, freq - 1)
    t = f - i
    s = t * t * (3.0 - 2.0 * t)
    out = 0.0
    for dx in (0, 1):
        wx = s[:, 0] if dx else 1.0 - s[:, 0]
        for dy in (0, 1):
            wy = s[:, 1] if dy else 1.0 - s[:, 1]
            for dz in (0, 1):
                wz = s[:, 2] if dz else 1.0 - s[:, 2]
                out = out + wx * wy * wz * lat[i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz]
    return out


def fbm_3d(p, octaves=4, seed=0, base_freq=4):
    rng = np.random.default_rng(seed)
    out = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, float(base_freq), 0.0
    for _ in range(octaves):
        out += amp * _value_noise_3d(p, int(round(freq)), rng)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return out / norm


# ----------------------------------------------------------------------------
# meshes
# ----------------------------------------------------------------------------


def heightfield(n: int = 1024, amplitude: float = 0.2, octaves: int = 6, lacunarity: float = 2.0,
                gain: float = 0.5, seed: int = 1) -> np.ndarray:
    """n x n cells of size 1.0 (extent n), two triangles per cell, 2*n*n triangles.

    Height = amplitude * n * fBm(x/n, z/n).  Triangle id = 2*(j*n + i) + k with
    k=0: (v00, v10, v11), k=1: (v00, v11, v01); both normals point down (-y).
    """
    g = np.arange(n + 1, dtype=np.float64)
    X, Z = np.meshgrid(g, g, indexing="xy")          # X[j, i] = i, Z[j, i] = j
    H = amplitude * n * fbm_2d(X / n, Z / n, octaves=octaves, lacunarity=lacunarity, gain=gain, seed=seed)
    P = np.stack([X, H, Z], axis=-1).astype(np.float32)  # [j, i, 3]
    v00 = P[:-1, :-1]
    v10 = P[:-1, 1:]
    v01 = P[1:, :-1]
    v11 = P[1:, 1:]
    t0 = np.concatenate([v00, v10, v11], axis=-1)
    t1 = np.concatenate([v00, v11, v01], axis=-1)
    tris = np.stack([t0, t1], axis=2)                 # [j, i, k, 9]
    return np.ascontiguousarray(tris.reshape(-1, 9))


def uv_sphere(center, radius, segments=16, rings=8) -> np.ndarray:
    """segments x rings bands, 2 triangles per quad including the degenerate pole
    triangles (a repeated pole vertex makes them always miss, Triangle.impl.h:292).
    16 x 8 gives 256 triangles.  Normals point inward."""
    c = np.asarray(center, dtype=np.float64)
    th = np.linspace(0.0, math.pi, rings + 1)            # polar
    ph = np.linspace(0.0, 2.0 * math.pi, segments + 1)   # azimuth
    V = np.zeros((rings + 1, segments + 1, 3))
    V[..., 0] = np.sin(th)[:, None] * np.cos(ph)[None, :]
    V[..., 1] = np.cos(th)[:, None] * np.ones_like(ph)[None, :]
    V[..., 2] = np.sin(th)[:, None] * np.sin(ph)[None, :]
    V[0, :] = (0.0, 1.0, 0.0)
    V[-1, :] = (0.0, -1.0, 0.0)
    V[:, -1] = V[:, 0]
    V = (c + radius * V).astype(np.float32)
    out = []
    for r in range(rings):
        for s in range(segments):
            p00, p01 = V[r, s], V[r, s + 1]
            p10, p11 = V[r + 1, s], V[r + 1, s + 1]
            out.append(np.concatenate([p00, p01, p11]))
            out.append(np.concatenate([p00, p11, p10]))
    tris = np.array(out, dtype=np.float32)
    return _orient_inward(tris, c)


def icosphere(center, radius, level: int, displace: float = 0.0, seed: int = 0) -> np.ndarray:
    """Subdivided icosahedron, 20*4^level triangles, optionally displaced radially by
    displace*radius*fBm3D(direction).  Normals point inward."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                      [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                      [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                      [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        nv = verts.shape[0]
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]], axis=0)
        e = np.sort(e, axis=1)
        key = e[:, 0] * (nv + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (nv + 1)
        b = uniq % (nv + 1)
        mid = verts[a] + verts[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        verts = np.concatenate([verts, mid], axis=0)
        nf = faces.shape[0]
        m01 = nv + inv[:nf]
        m12 = nv + inv[nf:2 * nf]
        m20 = nv + inv[2 * nf:]
        f0, f1, f2 = faces[:, 0], faces[:, 1], faces[:, 2]
        faces = np.concatenate([np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
                                np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=0)
    r = np.full(verts.shape[0], radius, dtype=np.float64)
    if displace:
        r = r * (1.0 + displace * fbm_3d(verts * 0.5 + 0.5, seed=seed))
    P = (np.asarray(center, np.float64) + verts * r[:, None]).astype(np.float32)
    tris = np.concatenate([P[faces[:, 0]], P[faces[:, 1]], P[faces[:, 2]]], axis=1).astype(np.float32)
    return _orient_inward(tris, np.asarray(center, np.float64))


def _orient_inward(tris: np.ndarray, center) -> np.ndarray:
    a = tris[:, 0:3].astype(np.float64)
    b = tris[:, 3:6].astype(np.float64)
    c = tris[:, 6:9].astype(np.float64)
    n = np.cross(b - a, c - a)
    cen = (a + b + c) / 3.0
    flip = np.einsum("ij,ij->i", n, cen - center) > 0    # outward -> swap b, c
    out = tris.copy()
    out[flip, 3:6] = tris[flip, 6:9]
    out[flip, 6:9] = tris[flip, 3:6]
    return out


def ground_quad(y: float, half: float, cx: float = 0.0, cz: float = 0.0) -> np.ndarray:
    """Two triangles at height y; normals point down so rays from above see front faces."""
    a = (cx - half, y, cz - half)
    b = (cx + half, y, cz - half)
    c = (cx + half, y, cz + half)
    d = (cx - half, y, cz + half)
    return np.array([a + b + c, a + c + d], dtype=np.float32)


# ----------------------------------------------------------------------------
# ray batches
# ----------------------------------------------------------------------------


def scene_aabb(tris: np.ndarray):
    p = tris.reshape(-1, 3)
    return p.min(axis=0), p.max(axis=0)


def incoherent_rays(tris: np.ndarray, n: int, seed: int = 7) -> np.ndarray:
    """n rays: origins uniform in the scene AABB, directions uniform on S^2
    (z = 1 - 2u, phi = 2 pi v).  Returns float32 (n, 6) = (o, d), d unit length."""
    lo, hi = scene_aabb(tris)
    rng = np.random.Generator(np.random.Philox(seed))
    o = lo + (hi - lo) * rng.random((n, 3))
    u = rng.random(n)
    v = rng.random(n)
    z = 1.0 - 2.0 * u
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * math.pi * v
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    return np.ascontiguousarray(np.concatenate([o, d], axis=1).astype(np.float32))


@dataclass
class Camera:
    """Pinhole camera = DefaultPerspectiveLense(fov) + Transform (position, axis vectors).

    DefaultPerspectiveLense.impl.cuh:5-15: focal = 0.5 / tan(fov / (2*RADIAN)), lens ray
    (sx, sy, focal) normalized; Transform::direction (Transform.impl.h:166-168) maps it to
    world as dir.x*right + dir.y*up + dir.z*forward, origin = position."""
    position: tuple
    right: tuple
    up: tuple
    forward: tuple
    fov_deg: float = 60.0

    @property
    def focal(self) -> float:
        radian = np.float32(self.fov_deg) / (np.float32(2) * np.float32(57.29577951308233))
        return float(np.float32(0.5) / np.float32(math.tan(float(radian))))

    @staticmethod
    def look_at(position, target, fov_deg=60.0, world_up=(0.0, 1.0, 0.0)):
        p = np.asarray(position, np.float64)
        f = np.asarray(target, np.float64) - p
        f /= np.linalg.norm(f)
        r = np.cross(np.asarray(world_up, np.float64), f)
        r /= np.linalg.norm(r)
        u = np.cross(f, r)
        f32 = lambda v: tuple(float(x) for x in np.asarray(v, np.float32))
        return Camera(f32(p), f32(r), f32(u), f32(f), float(fov_deg))

    def as_array(self) -> np.ndarray:
        """16 floats: position, right, up, forward, focal, 3 pad."""
        return np.array(list(self.position) + list(self.right) + list(self.up) + list(self.forward)
                        + [self.focal, 0.0, 0.0, 0.0], dtype=np.float32)


def primary_rays(cam: Camera, width: int, height: int) -> np.ndarray:
    """One pixel-centre ray per pixel in row-major (y, x) order, fp32 throughout, following
    DumbRenderer.cu:190-232 (HEIGHT boxing: pixelSize = 1/H, sample at
    ((x - W/2) + 0.5, (H/2 - y) + 0.5) * pixelSize for 1x1 spp) and the pinhole lens."""
    f = np.float32
    ps = f(1.0) / f(height)
    xs = np.arange(width, dtype=np.float32)
    ys = np.arange(height, dtype=np.float32)
    px = (xs - f(width) / f(2.0)) * ps
    py = (f(height) / f(2.0) - ys) * ps
    sx = (px + (f(0.0) + f(0.5)) * ps)[None, :].repeat(height, 0)
    sy = (py + (f(0.0) + f(0.5)) * ps)[:, None].repeat(width, 1)
    sz = np.full_like(sx, f(cam.focal))
    mag = np.sqrt(sx * sx + sy * sy + sz * sz)
    inv = f(1.0) / mag
    dx, dy, dz = sx * inv, sy * inv, sz * inv
    R = np.asarray(cam.right, np.float32)
    U = np.asarray(cam.up, np.float32)
    F = np.asarray(cam.forward, np.float32)
    d = [(dx * R[k] + dy * U[k]) + dz * F[k] for k in range(3)]
    o = [np.full_like(dx, f(cam.position[k])) for k in range(3)]
    rays = np.stack(o + d, axis=-1).reshape(-1, 6)
    return np.ascontiguousarray(rays.astype(np.float32))


# ----------------------------------------------------------------------------
# named workloads
# ----------------------------------------------------------------------------


def raycast_workload(cells: int = 1024, n_incoherent: int = 1 << 22, width: int = 1920, height: int = 1080,
                     seed: int = 1):
    """Ray-cast microbenchmark: fBm heightfield (2*cells^2 tris) + coherent primaries
    (width x height) + n_incoherent rays.  Returns (tris, coherent_rays, incoherent_rays, camera)."""
    tris = heightfield(cells, seed=seed)
    lo, hi = scene_aabb(tris)
    ext = float(cells)
    cam = Camera.look_at((0.5 * ext, float(hi[1]) + 0.35 * ext, -0.25 * ext),
                         (0.5 * ext, float(lo[1]), 0.55 * ext), fov_deg=60.0)
    coherent = primary_rays(cam, width, height) if width and height else np.zeros((0, 6), np.float32)
    incoherent = incoherent_rays(tris, n_incoherent, seed=seed + 6)
    return tris, coherent, incoherent, cam


@dataclass
class Scene:
    tris: np.ndarray                    # (n, 9) float32
    mat_id: np.ndarray                  # (n,) int32
    materials: list = field(default_factory=list)   # list of dicts
    lights: list = field(default_factory=list)      # list of dicts (area quads)
    camera: Camera | None = None


def spheres_scene(n_spheres: int = 8, seed: int = 0, n_lights: int = 1, ground_material: bool = True) -> Scene:
    """Ground quad (2 tris) + n UV spheres (16x8, 256 tris each), centres U[-4,4]^3,
    radii U(0.3,1.0), Lambert albedo U(0.2,0.8)^3 per sphere; area-light quads
    1x1 at y=5 with emission 10.  ground_material=False gives the ground sphere 0's
    material, so there is exactly one material per sphere."""
    rng = np.random.Generator(np.random.Philox(seed))
    parts = [ground_quad(-4.0 - 1.0, 12.0)]
    mids = [np.zeros(2, np.int32)]
    mats = [dict(kind="lambert", albedo=(0.5, 0.5, 0.5))] if ground_material else []
    for s in range(n_spheres):
        c = rng.uniform(-4.0, 4.0, 3)
        r = rng.uniform(0.3, 1.0)
        alb = rng.uniform(0.2, 0.8, 3).astype(np.float32)
        parts.append(uv_sphere(c, r))
        mid = len(mats)
        mats.append(dict(kind="lambert", albedo=tuple(float(x) for x in alb)))
        mids.append(np.full(256, mid, np.int32))
    lights = []
    for k in range(n_lights):
        if n_lights == 1:
            cx, cz = 0.0, 0.0
        else:
            ang = 2.0 * math.pi * k / n_lights
            cx, cz = 3.0 * math.cos(ang), 3.0 * math.sin(ang)
        lights.append(dict(kind="quad", corner=(cx - 0.5, 5.0, cz - 0.5), edge_u=(1.0, 0.0, 0.0),
                           edge_v=(0.0, 0.0, 1.0), emission=(10.0, 10.0, 10.0)))
    cam = Camera.look_at((0.0, 2.0, -14.0), (0.0, -1.0, 0.0), fov_deg=60.0)
    return Scene(np.ascontiguousarray(np.concatenate(parts, 0)), np.concatenate(mids).astype(np.int32),
                 mats, lights, cam)


def icosphere_scene(level: int = 7, seed: int = 3) -> Scene:
    """3 displaced icospheres (Lambert, mirror, Schlick metal F0=(1.0,0.78,0.34)) on a
    ground quad, two area lights."""
    rng = np.random.Generator(np.random.Philox(seed))
    centers = [(-2.4, 0.0, 0.0), (0.0, 0.0, 1.2), (2.4, 0.0, 0.0)]
    parts = [ground_quad(-1.0, 10.0)]
    mids = [np.zeros(2, np.int32)]
    mats = [dict(kind="lambert", albedo=(0.6, 0.6, 0.6)),
            dict(kind="lambert", albedo=tuple(float(x) for x in rng.uniform(0.2, 0.8, 3).astype(np.float32))),
            dict(kind="mirror", albedo=(1.0, 1.0, 1.0)),
            dict(kind="metal", albedo=(1.0, 0.78, 0.34))]
    for k, c in enumerate(centers):
        t = icosphere(c, 1.0, level, displace=0.05, seed=seed + k)
        parts.append(t)
        mids.append(np.full(t.shape[0], k + 1, np.int32))
    lights = [dict(kind="quad", corner=(-3.0, 5.0, -2.0), edge_u=(1.5, 0.0, 0.0), edge_v=(0.0, 0.0, 1.5),
                   emission=(12.0, 11.0, 10.0)),
              dict(kind="quad", corner=(2.0, 4.0, -3.0), edge_u=(1.0, 0.0, 0.0), edge_v=(0.0, 0.0, 1.0),
                   emission=(6.0, 7.0, 9.0))]
    cam = Camera.look_at((0.0, 2.5, -7.5), (0.0, 0.0, 0.3), fov_deg=60.0)
    return Scene(np.ascontiguousarray(np.concatenate(parts, 0)), np.concatenate(mids).astype(np.int32),
                 mats, lights, cam)


def batch_workload(n_requests: int = 64, n_scenes: int = 8, seed: int = 2024):
    """Batched uneven requests: n_scenes seeded scenes built like icosphere_scene at subdivision
    levels 5-7 (61,442-983,042 tris each), and n_requests requests over them (request i uses scene
    i % n_scenes) with resolution from {128^2, 256^2, 512^2, 1024^2}, samples per pixel from
    {1, 4, 16, 64}, max depth from {2, 4, 8}, Russian roulette from depth 3 (survival 0.75).
    Returns (scenes, requests); a request is a dict of scene, width, height, fsaa, depth, seed and
    request (the Philox counter word)."""
    rng = np.random.default_rng(seed)
    scenes = [icosphere_scene(level=int(rng.integers(5, 8)), seed=100 + k) for k in range(n_scenes)]
    reqs = []
    for i in range(n_requests):
        res = int([128, 256, 512, 1024][rng.integers(0, 4)])
        f = int([1, 2, 4, 8][rng.integers(0, 4)])
        depth = int([2, 4, 8][rng.integers(0, 3)])
        reqs.append(dict(scene=i % n_scenes, width=res, height=res, fsaa=(f, f), depth=depth, seed=7000 + i,
                         request=i + 1))
    return scenes, reqs


def random_soup(n: int, seed: int = 0, extent: float = 10.0, size: float = 1.0) -> np.ndarray:
    """Random triangle soup for parity tests (float32)."""
    rng = np.random.Generator(np.random.Philox(seed))
    c = rng.uniform(-extent, extent, (n, 1, 3))
    v = c + rng.uniform(-size, size, (n, 3, 3))
    return np.ascontiguousarray(v.reshape(n, 9).astype(np.float32))
