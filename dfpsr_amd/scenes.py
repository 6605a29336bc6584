"""Seeded synthetic scenes for the polygon renderer.

Every scene is what one renderer_begin / renderer_giveTask(model)* /
renderer_end frame of the reference consumes (Source/SDK/terrain/main.cpp:381-421):
a model of points and parts (Model.h:39-133), a camera made by
Camera::createPerspective (Camera.h:144-149) and two canvases.

All randomness is SplitMix64 seeded per scene, and floats are made as
``(x >> 40) * 2**-24`` so that the same bits come out on every platform.
Scene geometry is produced in float32 exactly once here; the CPU oracle, the
reference driver and the GPU library all consume these same arrays.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs for `seed` (uint64)."""
    with np.errstate(over="ignore"):
        state = np.uint64(seed & _M64) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = state
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


class Rng:
    """Counter-based SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = seed & _M64
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        out = splitmix64((self.seed + self.pos * 0x9E3779B97F4A7C15) & _M64, n)
        self.pos += n
        return out

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        """Float32 in [lo, hi): (x >> 40) * 2^-24, then lo + (hi - lo) * f in float32."""
        f = (self.u64(n) >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
        return (np.float32(lo) + np.float32(hi - lo) * f).astype(np.float32)

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """Integers in [lo, hi)."""
        span = hi - lo
        return (lo + (self.u64(n) % np.uint64(span)).astype(np.int64)).astype(np.int64)


@dataclass
class Texture:
    """Level-0 pixels of a power-of-two RGBA8 texture (packed R | G<<8 | B<<16 | A<<24)
    and the number of resolutions to give texture_create_RgbaU8 (textureAPI.cpp:65-78)."""
    width: int
    height: int
    resolutions: int
    level0: np.ndarray  # (height, width) uint32


@dataclass
class Part:
    polygon_start: int
    polygon_end: int
    diffuse: int = -1  # texture index or -1
    light: int = -1


@dataclass
class Scene:
    name: str
    width: int
    height: int
    points: np.ndarray            # (P, 3) float32
    indices: np.ndarray           # (M, 4) int32, [:, 3] = -1 for triangles
    colors: np.ndarray            # (M, 4, 4) float32 RGBA per corner
    texcoords: np.ndarray         # (M, 4, 4) float32 (u1, v1, u2, v2) per corner
    parts: List[Part]
    textures: List[Texture]
    cam_location: np.ndarray      # (12,) float32: position, xAxis, yAxis, zAxis
    width_slope: float
    near_clip: float = 0.01
    far_clip: float = 1000.0
    perspective: bool = True
    filter: int = 0               # 0 Solid, 1 Alpha
    instances: np.ndarray = field(default_factory=lambda: identity_transform()[None, :])
    has_color: bool = True
    has_depth: bool = True
    initial_color: Optional[np.ndarray] = None   # (H, W) uint32
    initial_depth: Optional[np.ndarray] = None   # (H, W) float32
    seed: int = 0
    # Optional (M, 3) records of the caller's ProjectedPoints (cs, is, flat) per triangle corner:
    # the frame then gives the triangles with renderer_giveTask_triangle instead of the model.
    projected: Optional[np.ndarray] = None
    # Optional occluder calls of the renderer around the instances, in order: tuples (kind,
    # instance the call comes before, box minimum, box maximum, box transform) with kind one of
    # "box" (renderer_occludeFromBox), "top_rows", "existing" (renderer_occludeFrom...) and
    # "query" (renderer_isBoxVisible); the box fields are unused by top_rows and existing.
    occluders: Optional[list] = None

    @property
    def triangle_count(self) -> int:
        return int(np.sum(np.where(self.indices[:, 3] >= 0, 2, 1)))

    @property
    def polygon_count(self) -> int:
        return int(self.indices.shape[0])


def identity_transform() -> np.ndarray:
    return np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float32)


def look_transform(position, yaw: float, pitch: float) -> np.ndarray:
    """Camera location: +z forward, +y up, +x right, rotated by yaw (about y) then
    pitch (positive looks down).  Axes are computed in float64 and stored as float32."""
    cy, sy = math.cos(yaw), math.sin(yaw)
    cp, sp = math.cos(pitch), math.sin(pitch)
    forward = np.array([sy * cp, -sp, cy * cp])
    right = np.array([cy, 0.0, -sy])
    up = np.cross(forward, right)
    out = np.concatenate([np.asarray(position, dtype=np.float64), right, up, forward])
    return out.astype(np.float32)


def pack_rgba(r, g, b, a) -> np.ndarray:
    r, g, b, a = (np.asarray(c).astype(np.uint32) for c in (r, g, b, a))
    return (r | (g << np.uint32(8)) | (b << np.uint32(16)) | (a << np.uint32(24))).astype(np.uint32)


# ---------------------------------------------------------------------------
# Config 0: 640x480, 2000 vertex-coloured triangles in a [-1,1]^3 cube.
# ---------------------------------------------------------------------------

def cube_soup(width=640, height=480, triangles=2000, seed=0, spread=0.1, cam_distance=2.5,
              width_slope=1.0, far_clip=100.0, alpha=1.0, name="cube_soup") -> Scene:
    rng = Rng(seed)
    centers = rng.uniform(triangles * 3, -1.0, 1.0).reshape(triangles, 3)
    offsets = rng.uniform(triangles * 9, -spread, spread).reshape(triangles, 3, 3)
    pts = (centers[:, None, :] + offsets).astype(np.float32).reshape(-1, 3)
    rgb = rng.integers(triangles * 9, 0, 256).reshape(triangles, 3, 3).astype(np.float32) / np.float32(255.0)
    colors = np.ones((triangles, 4, 4), dtype=np.float32)
    colors[:, :3, :3] = rgb
    colors[:, :, 3] = np.float32(alpha)
    idx = np.full((triangles, 4), -1, dtype=np.int32)
    idx[:, :3] = np.arange(triangles * 3, dtype=np.int32).reshape(triangles, 3)
    tex = np.zeros((triangles, 4, 4), dtype=np.float32)
    cam = identity_transform()
    cam[2] = np.float32(-cam_distance)
    return Scene(name=name, width=width, height=height, points=pts, indices=idx, colors=colors,
                 texcoords=tex, parts=[Part(0, triangles)], textures=[], cam_location=cam,
                 width_slope=width_slope, far_clip=far_clip, seed=seed)


# ---------------------------------------------------------------------------
# Config 1: 1920x1080 textured terrain.
# ---------------------------------------------------------------------------

def value_noise(rng_seed: int, lattice: int, xs: np.ndarray, zs: np.ndarray) -> np.ndarray:
    """2D value noise on a periodic lattice of lattice x lattice random values in [0,1),
    smoothstep-interpolated.  xs, zs are lattice coordinates (float64)."""
    vals = Rng(rng_seed).uniform(lattice * lattice).astype(np.float64).reshape(lattice, lattice)
    x0 = np.floor(xs).astype(np.int64)
    z0 = np.floor(zs).astype(np.int64)
    fx = xs - x0
    fz = zs - z0
    sx = fx * fx * (3 - 2 * fx)
    sz = fz * fz * (3 - 2 * fz)

    def v(ix, iz):
        return vals[iz % lattice, ix % lattice]

    a = v(x0, z0) * (1 - sx) + v(x0 + 1, z0) * sx
    b = v(x0, z0 + 1) * (1 - sx) + v(x0 + 1, z0 + 1) * sx
    return a * (1 - sz) + b * sz


def terrain_heights(n: int, seed: int, octaves=6, amplitude=40.0, lacunarity=2.0, gain=0.5,
                    base_period=64.0, lattice=256) -> np.ndarray:
    """fBm of value noise in [0,1), centred and scaled so heights span about +-amplitude/2."""
    zs, xs = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    total = np.zeros((n, n))
    norm = 0.0
    amp = 1.0
    freq = 1.0 / base_period
    for o in range(octaves):
        total += amp * value_noise(seed * 1000 + o + 1, lattice, xs * freq, zs * freq)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return ((total / norm - 0.5) * amplitude).astype(np.float32)


def noise_checker_texture(size: int, seed: int, resolutions: int, check=64) -> Texture:
    rng = Rng(seed)
    noise = rng.integers(size * size * 3, 96, 256).reshape(size, size, 3)
    yy, xx = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    checker = (((xx // check) + (yy // check)) % 2).astype(np.int64)
    scale = np.where(checker == 1, 256, 160)[..., None]
    rgb = (noise * scale) >> 8
    px = pack_rgba(rgb[..., 0], rgb[..., 1], rgb[..., 2], np.full((size, size), 255))
    return Texture(size, size, resolutions, px)


def terrain_mesh(n: int, heights: np.ndarray, uv_repeat: float, light=(0.3, 1.0, 0.5), x0=None, z0=None):
    """Quad grid of n x n points with spacing 1, centred on the origin; quads alternate their
    diagonal like Source/SDK/terrain/main.cpp:120-176.  Vertex colour is a per-point Lambert term."""
    if x0 is None:
        x0 = -(n - 1) / 2.0
    if z0 is None:
        z0 = -(n - 1) / 2.0
    zs, xs = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    pts = np.stack([xs + x0, heights.astype(np.float64), zs + z0], axis=-1).astype(np.float32).reshape(-1, 3)
    h = heights.astype(np.float64)
    hp = np.pad(h, 1, mode="edge")
    dhdx = (hp[1:-1, 2:] - hp[1:-1, :-2]) * 0.5
    dhdz = (hp[2:, 1:-1] - hp[:-2, 1:-1]) * 0.5
    nrm = np.stack([-dhdx, np.ones_like(h), -dhdz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    l = np.asarray(light, dtype=np.float64)
    l /= np.linalg.norm(l)
    lambert = np.clip(nrm @ l, 0.0, 1.0) * 0.8 + 0.2
    shade = lambert.astype(np.float32).reshape(-1)

    qz, qx = np.meshgrid(np.arange(1, n), np.arange(1, n), indexing="ij")
    qx = qx.reshape(-1)
    qz = qz.reshape(-1)
    px, pz = qx - 1, qz - 1
    ia = px + pz * n
    ib = qx + pz * n
    ic = qx + qz * n
    id_ = px + qz * n
    even = ((qx + qz) % 2) == 0
    # Looking down from +y with +z forward on screen, (A, D, C, B) runs clockwise.
    quad = np.where(even[:, None], np.stack([ia, id_, ic, ib], 1), np.stack([id_, ic, ib, ia], 1)).astype(np.int32)
    m = quad.shape[0]
    scale = np.float32(uv_repeat / (n - 1))
    grid_x = np.tile(np.arange(n, dtype=np.float32), n)
    grid_z = np.repeat(np.arange(n, dtype=np.float32), n)
    u = grid_x * scale
    v = grid_z * scale
    tex = np.zeros((m, 4, 4), dtype=np.float32)
    tex[:, :, 0] = u[quad]
    tex[:, :, 1] = v[quad]
    colors = np.ones((m, 4, 4), dtype=np.float32)
    colors[:, :, 0] = shade[quad]
    colors[:, :, 1] = shade[quad]
    colors[:, :, 2] = shade[quad]
    return pts, quad, colors, tex


def terrain(width=1920, height=1080, grid=256, seed=1, texture_size=1024, resolutions=11,
            uv_repeat=4.0, fov_deg=60.0, cam_height=30.0, pitch_deg=35.0, cam_back=10.0,
            name="terrain") -> Scene:
    heights = terrain_heights(grid, seed)
    pts, quad, colors, tex = terrain_mesh(grid, heights, uv_repeat)
    texture = noise_checker_texture(texture_size, seed + 17, resolutions)
    half = (grid - 1) / 2.0
    cam = look_transform((0.0, cam_height, -half - cam_back), 0.0, math.radians(pitch_deg))
    return Scene(name=name, width=width, height=height, points=pts, indices=quad, colors=colors,
                 texcoords=tex, parts=[Part(0, quad.shape[0], diffuse=0)], textures=[texture],
                 cam_location=cam, width_slope=float(np.float32(math.tan(math.radians(fov_deg) / 2))),
                 near_clip=0.01, far_clip=1000.0, seed=seed)


# ---------------------------------------------------------------------------
# Small hand-made scenes for edge cases.
# ---------------------------------------------------------------------------

def single_triangles(width, height, tris_screen, colors=None, depth=2.0, textured=False, seed=5,
                     name="micro") -> Scene:
    """Triangles given in pixel coordinates at a common camera-space depth; unprojected through a
    width_slope = 1 camera at the origin so they land (up to rounding) where asked."""
    tris_screen = np.asarray(tris_screen, dtype=np.float64).reshape(-1, 3, 2)
    n = tris_screen.shape[0]
    depths = np.broadcast_to(np.asarray(depth, dtype=np.float64), (n,)) if np.ndim(depth) else np.full(n, depth)
    height_slope = height / width
    pts = []
    for t in range(n):
        z = depths[t]
        for (sx, sy) in tris_screen[t]:
            x = (sx / width - 0.5) * 2.0 * z
            y = -(sy / height - 0.5) * 2.0 * z * height_slope
            pts.append((x, y, z))
    pts = np.asarray(pts, dtype=np.float32)
    idx = np.full((n, 4), -1, dtype=np.int32)
    idx[:, :3] = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    rng = Rng(seed)
    cols = np.ones((n, 4, 4), dtype=np.float32)
    if colors is None:
        cols[:, :3, :3] = rng.uniform(n * 9).reshape(n, 3, 3)
    else:
        cols[:, :3, :] = np.broadcast_to(np.asarray(colors, dtype=np.float32).reshape(n, -1, 4), (n, 3, 4))
    tex = np.zeros((n, 4, 4), dtype=np.float32)
    textures = []
    parts = [Part(0, n)]
    if textured:
        tex[:, :3, :2] = rng.uniform(n * 6, -0.5, 1.5).reshape(n, 3, 2)
        textures = [noise_checker_texture(64, seed + 3, 7, check=8)]
        parts = [Part(0, n, diffuse=0)]
    return Scene(name=name, width=width, height=height, points=pts, indices=idx, colors=cols,
                 texcoords=tex, parts=parts, textures=textures, cam_location=identity_transform(),
                 width_slope=1.0, far_clip=100.0, seed=seed)


# ---------------------------------------------------------------------------
# Config 2: thousands of independent jobs with uneven canvases and triangle counts.
# ---------------------------------------------------------------------------

def orbit_transform(distance: float, yaw: float, pitch: float) -> np.ndarray:
    """Camera on a sphere of `distance` around the origin, looking at it."""
    cy, sy = math.cos(yaw), math.sin(yaw)
    cp, sp = math.cos(pitch), math.sin(pitch)
    forward = np.array([-sy * cp, -sp, -cy * cp])
    position = -forward * distance
    return look_transform(position, math.atan2(forward[0], forward[2]), math.asin(-forward[1]))


def batch_jobs(count: int = 4096, seed: int = 2, min_tris: int = 100, max_tris: int = 20000,
               texture_count: int = 16) -> List[Scene]:
    """Uneven jobs: canvas width U[64,640] and height U[48,480] rounded to multiples of 4;
    triangle count log-uniform in [min_tris, max_tris]; half are random-cube soups (as config 0)
    and half terrain patches of n x n points with 2(n-1)^2 close to the drawn count (a 32x32
    patch is 1,922 triangles); half of the jobs sample a 256x256 RGBA8 texture with a full mip
    chain (one of `texture_count` shared textures); each camera orbits its object."""
    rng = Rng(seed)
    textures = [noise_checker_texture(256, seed * 100 + t, 9, check=16 + 8 * (t % 4)) for t in range(texture_count)]
    jobs = []
    for j in range(count):
        w = int(rng.integers(1, 64, 641)[0]) // 4 * 4
        h = int(rng.integers(1, 48, 481)[0]) // 4 * 4
        tris = int(round(math.exp(float(rng.uniform(1, math.log(min_tris), math.log(max_tris))[0]))))
        terrain_job = bool(rng.integers(1, 0, 2)[0])
        textured = bool(rng.integers(1, 0, 2)[0])
        tex_id = int(rng.integers(1, 0, texture_count)[0])
        yaw = float(rng.uniform(1, 0.0, 2.0 * math.pi)[0])
        pitch = float(rng.uniform(1, 0.15, 1.2)[0])
        job_seed = seed * 1000003 + j
        if terrain_job:
            n = max(4, int(round(math.sqrt(tris / 2.0))) + 1)
            heights = terrain_heights(n, job_seed, octaves=4, amplitude=0.3 * n, base_period=max(4.0, n / 2.0), lattice=32)
            pts, quad, colors, tex = terrain_mesh(n, heights, uv_repeat=2.0)
            sc = Scene(name=f"job{j}", width=w, height=h, points=pts, indices=quad, colors=colors, texcoords=tex,
                       parts=[Part(0, quad.shape[0], diffuse=0 if textured else -1)],
                       textures=[textures[tex_id]] if textured else [],
                       cam_location=orbit_transform(0.9 * n, yaw, pitch), width_slope=1.0, seed=job_seed)
        else:
            sc = cube_soup(width=w, height=h, triangles=tris, seed=job_seed, name=f"job{j}")
            sc.cam_location = orbit_transform(2.5, yaw, pitch * 0.5)
            if textured:
                r2 = Rng(job_seed + 77)
                sc.texcoords[:, :3, :2] = r2.uniform(tris * 6, -1.0, 2.0).reshape(tris, 3, 2)
                sc.parts = [Part(0, tris, diffuse=0)]
                sc.textures = [textures[tex_id]]
        sc.texture_ids = [tex_id] if textured else []
        jobs.append(sc)
    return jobs


# ---------------------------------------------------------------------------
# Configs 3 and 4: a displaced geodesic sphere seen from many views.
# ---------------------------------------------------------------------------

def look_at(position, target=(0.0, 0.0, 0.0), world_up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Camera location at `position` looking at `target`: x right, y up, z forward."""
    p = np.asarray(position, dtype=np.float64)
    forward = np.asarray(target, dtype=np.float64) - p
    forward /= np.linalg.norm(forward)
    up = np.asarray(world_up, dtype=np.float64)
    if abs(np.dot(up, forward)) > 0.999:
        up = np.array([0.0, 0.0, 1.0])
    right = np.cross(up, forward)
    right /= np.linalg.norm(right)
    up = np.cross(forward, right)
    return np.concatenate([p, right, up, forward]).astype(np.float32)


def fibonacci_directions(count: int, seed: int = 0) -> np.ndarray:
    """`count` near-uniform unit directions (Fibonacci sphere), rotated by a seeded angle."""
    k = np.arange(count, dtype=np.float64) + 0.5
    y = 1.0 - 2.0 * k / count
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    phi = k * math.pi * (3.0 - math.sqrt(5.0)) + float(Rng(seed).uniform(1, 0.0, 2.0 * math.pi)[0])
    return np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)


def geodesic_sphere(frequency: int, seed: int, amplitude: float = 0.1, waves: int = 12):
    """Icosahedron whose faces are split into frequency^2 triangles (20 * frequency^2 in all),
    pushed to the unit sphere and displaced radially by amplitude * a seeded sum of `waves`
    random plane waves.  Face-edge vertices are duplicated per face (same position).
    Returns points (P, 3), triangles (T, 3) and per-corner uv (T, 3, 2)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                      [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    faces = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
             [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
             [8, 6, 7], [9, 8, 1]]
    n = frequency
    ii, jj = [], []
    for i in range(n + 1):
        for j in range(n + 1 - i):
            ii.append(i)
            jj.append(j)
    ii = np.array(ii)
    jj = np.array(jj)
    index = -np.ones((n + 1, n + 1), dtype=np.int64)
    index[ii, jj] = np.arange(len(ii))
    tri_local = []
    for i in range(n):
        for j in range(n - i):
            tri_local.append((index[i, j], index[i + 1, j], index[i, j + 1]))
            if i + j < n - 1:
                tri_local.append((index[i + 1, j], index[i + 1, j + 1], index[i, j + 1]))
    tri_local = np.array(tri_local, dtype=np.int64)
    pts, tris = [], []
    for f, (a, b, c) in enumerate(faces):
        A, B, C = verts[a], verts[b], verts[c]
        p = A[None, :] + (B - A)[None, :] * (ii / n)[:, None] + (C - A)[None, :] * (jj / n)[:, None]
        tris.append(tri_local + f * len(ii))
        pts.append(p)
    pts = np.concatenate(pts)
    tris = np.concatenate(tris)
    d = pts / np.linalg.norm(pts, axis=1, keepdims=True)
    rng = Rng(seed)
    k = rng.uniform(waves * 3, -1.0, 1.0).reshape(waves, 3).astype(np.float64) * 6.0
    phase = rng.uniform(waves, 0.0, 2.0 * math.pi).astype(np.float64)
    noise = np.sin(d @ k.T + phase[None, :]).sum(axis=1) / math.sqrt(waves)
    noise /= max(1e-9, np.abs(noise).max())
    pts = d * (1.0 + amplitude * noise)[:, None]
    u = np.arctan2(d[:, 0], d[:, 2]) / (2.0 * math.pi) + 0.5
    v = np.arccos(np.clip(d[:, 1], -1.0, 1.0)) / math.pi
    uv = np.stack([u[tris], v[tris]], axis=-1)
    # Triangles across the u seam: lift the small u values by one lap (the texture repeats).
    span = uv[:, :, 0].max(axis=1) - uv[:, :, 0].min(axis=1)
    seam = span > 0.5
    uv[:, :, 0] = np.where(seam[:, None] & (uv[:, :, 0] < 0.5), uv[:, :, 0] + 1.0, uv[:, :, 0])
    # The icosahedron's faces wind counter-clockwise seen from outside, which the screen's
    # downward y turns into the clockwise order ITriangle2D::isFrontfacing keeps.
    return pts.astype(np.float32), tris.astype(np.int32), uv.astype(np.float32)


def sphere_views(frequency: int, views: int, width: int, height: int, seed: int = 3, radius: float = 3.0,
                 fov_deg: float = 60.0, name: str = "sphere") -> List[Scene]:
    """One Scene per view of the displaced geodesic sphere, cameras on a Fibonacci sphere."""
    pts, tris, uv = geodesic_sphere(frequency, seed)
    m = tris.shape[0]
    idx = np.full((m, 4), -1, dtype=np.int32)
    idx[:, :3] = tris
    tex = np.zeros((m, 4, 4), dtype=np.float32)
    tex[:, :3, :2] = uv
    colors = np.ones((m, 4, 4), dtype=np.float32)
    dirs = fibonacci_directions(views, seed)
    out = []
    for k in range(views):
        cam = look_at(dirs[k] * radius)
        out.append(Scene(name=f"{name}_view{k}", width=width, height=height, points=pts, indices=idx, colors=colors,
                         texcoords=tex, parts=[Part(0, m)], textures=[], cam_location=cam,
                         width_slope=float(np.float32(math.tan(math.radians(fov_deg) / 2))), near_clip=0.01,
                         far_clip=100.0, seed=seed))
    return out


def float_texture(size: int, seed: int, smooth: int = 8) -> np.ndarray:
    """A seeded RGBA float32 texture (size x size x 4) in [0, 1): blurred noise plus stripes."""
    rng = Rng(seed)
    base = rng.uniform((size // smooth) * (size // smooth) * 4).reshape(size // smooth, size // smooth, 4)
    up = np.repeat(np.repeat(base, smooth, axis=0), smooth, axis=1)
    yy, xx = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    stripes = 0.5 + 0.5 * np.sin(2 * math.pi * (xx + 2 * yy) / (size / 8.0))
    out = 0.6 * up + 0.4 * stripes[..., None]
    out[..., 3] = 1.0
    return out.astype(np.float32)
