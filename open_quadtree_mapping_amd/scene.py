"""Seeded synthetic planar scenes for the per-keyframe depth pipeline.

The reference ships no data (its only exercise is a rosbag that has to be
downloaded), so inputs are generated here: a handful of textured rectangles at
1-10 m, rendered with a pinhole camera into uint8 grayscale frames together with
the exact Euclidean depth (the depth convention of the reference README, the
distance along the ray).

Everything is a pure function of its seed (numpy PCG64 streams and an integer
hash for the lattice values); there is no global RNG state.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .se3 import se3_from_rt, se3_inv, rot_xyz

__all__ = ["Camera", "Plane", "Scene", "make_scene", "render", "lateral_trajectory",
           "forward_lateral_trajectory", "camera_for"]


@dataclass(frozen=True)
class Camera:
    W: int
    H: int
    fx: float
    fy: float
    cx: float
    cy: float


def camera_for(W: int, H: int) -> Camera:
    """Pinhole intrinsics scaled from the 752x480 f=365 camera."""
    s = W / 752.0
    return Camera(W, H, 365.0 * s, 365.0 * s, W / 2.0, H / 2.0)


@dataclass
class Plane:
    origin: np.ndarray      # (3,) world point at the plane centre
    e1: np.ndarray          # (3,) unit in-plane axis
    e2: np.ndarray          # (3,) unit in-plane axis
    normal: np.ndarray      # (3,) unit normal
    half: tuple             # half extents along e1, e2 (m)
    salt: int               # texture seed


@dataclass
class Scene:
    planes: list = field(default_factory=list)
    lattice: float = 0.08   # base lattice pitch of the value noise (m)
    octaves: int = 5


_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xC2B2AE3D27D4EB4F)
_M3 = np.uint64(0xBF58476D1CE4E5B9)
_M4 = np.uint64(0x94D049BB133111EB)


def _hash01(ix: np.ndarray, iy: np.ndarray, salt: int) -> np.ndarray:
    """splitmix64 finaliser of the lattice coordinates -> uniform [0, 1)."""
    with np.errstate(over="ignore"):
        h = (ix.astype(np.int64).view(np.uint64) * _M1) ^ (iy.astype(np.int64).view(np.uint64) * _M2)
        h = h ^ np.uint64(salt & 0xFFFFFFFFFFFFFFFF)
        h = (h ^ (h >> np.uint64(30))) * _M3
        h = (h ^ (h >> np.uint64(27))) * _M4
        h = h ^ (h >> np.uint64(31))
    return (h >> np.uint64(40)).astype(np.float64) / float(1 << 24)


def _value_noise(u: np.ndarray, v: np.ndarray, salt: int, octaves: int) -> np.ndarray:
    """Fractal value noise (lacunarity 2, gain 0.5), normalised to [0, 1]."""
    total = np.zeros_like(u)
    amp, norm, freq = 1.0, 0.0, 1.0
    for o in range(octaves):
        x = u * freq
        y = v * freq
        ix = np.floor(x)
        iy = np.floor(y)
        fx = x - ix
        fy = y - iy
        sx = fx * fx * (3.0 - 2.0 * fx)
        sy = fy * fy * (3.0 - 2.0 * fy)
        ix = ix.astype(np.int64)
        iy = iy.astype(np.int64)
        s = salt * 131 + o * 7919
        v00 = _hash01(ix, iy, s)
        v10 = _hash01(ix + 1, iy, s)
        v01 = _hash01(ix, iy + 1, s)
        v11 = _hash01(ix + 1, iy + 1, s)
        a = v00 + (v10 - v00) * sx
        b = v01 + (v11 - v01) * sx
        total += amp * (a + (b - a) * sy)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def make_scene(seed: int, n_planes: int = 6) -> Scene:
    """n_planes textured rectangles in front of a camera at the origin looking down +z.

    Centre depths ~ U(1, 10) m, normals within 30 deg of the optical axis,
    extents covering 25-60% of the view; the farthest one is enlarged so every
    ray of the trajectories below hits a surface.
    """
    rng = np.random.default_rng(seed)
    planes = []
    depths = np.sort(rng.uniform(1.0, 10.0, size=n_planes))
    for k in range(n_planes):
        z = depths[k]
        # centre somewhere inside the initial view (tan of half fov ~ 1)
        cxy = rng.uniform(-0.6, 0.6, size=2) * z
        tilt = np.deg2rad(rng.uniform(0.0, 30.0))
        az = rng.uniform(0.0, 2.0 * np.pi)
        n = np.array([np.sin(tilt) * np.cos(az), np.sin(tilt) * np.sin(az), -np.cos(tilt)])
        a = np.array([1.0, 0.0, 0.0])
        e1 = a - n * (a @ n)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        frac = rng.uniform(0.25, 0.60, size=2)
        half = (float(frac[0] * z), float(frac[1] * z * 0.7))
        planes.append(Plane(np.array([cxy[0], cxy[1], z]), e1, e2, n, half, int(rng.integers(1, 1 << 31))))
    far = planes[-1]
    far.origin = np.array([0.0, 0.0, depths[-1] + 2.0])
    far.half = (1e3, 1e3)
    return Scene(planes=planes)


def render(scene: Scene, cam: Camera, T_cw: np.ndarray):
    """Render one frame. T_cw (3x4) maps world points into the camera frame.

    Returns (image uint8 HxW, Euclidean depth float32 HxW; 0 where nothing is hit).
    """
    W, H = cam.W, cam.H
    R = T_cw[:, :3].astype(np.float64)
    t = T_cw[:, 3].astype(np.float64)
    o = -R.T @ t                                     # camera centre in world
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    rc = np.stack([(u - cam.cx) / cam.fx, (v - cam.cy) / cam.fy, np.ones_like(u)], axis=-1)
    dw = rc @ R                                      # R^T rc for every pixel
    best_t = np.full((H, W), np.inf)
    img = np.zeros((H, W), dtype=np.float64)
    for pl in scene.planes:
        denom = dw @ pl.normal
        with np.errstate(divide="ignore", invalid="ignore"):
            tt = ((pl.origin - o) @ pl.normal) / denom
        hit = np.isfinite(tt) & (tt > 1e-6) & (tt < best_t)
        if not hit.any():
            continue
        p = o + dw * tt[..., None]
        rel = p - pl.origin
        a = rel @ pl.e1
        b = rel @ pl.e2
        hit &= (np.abs(a) <= pl.half[0]) & (np.abs(b) <= pl.half[1])
        if not hit.any():
            continue
        val = _value_noise(a[hit] / scene.lattice, b[hit] / scene.lattice, pl.salt, scene.octaves)
        img[hit] = np.clip(np.round(40.0 + 175.0 * val), 0, 255)
        best_t[hit] = tt[hit]
    depth = np.where(np.isfinite(best_t), best_t * np.linalg.norm(rc, axis=-1), 0.0)
    return img.astype(np.uint8), depth.astype(np.float32)


def lateral_trajectory(n: int, step: float = 0.04, jitter_deg: float = 0.5, seed: int = 0):
    """n poses T_cw: +step metres along x per frame, per-frame absolute rotation
    jitter U(-jitter, +jitter) deg about each axis."""
    rng = np.random.default_rng(seed + 1009)
    out = []
    for k in range(n):
        ang = np.deg2rad(rng.uniform(-jitter_deg, jitter_deg, size=3))
        R_wc = rot_xyz(*ang)
        c = np.array([k * step, 0.0, 0.0])
        out.append(se3_inv(se3_from_rt(R_wc, c)))
    return out


def forward_lateral_trajectory(n: int, step: float = 0.05, seed: int = 0, jitter_deg: float = 0.0):
    """n poses T_cw moving step metres per frame along normalize(1, 0, 1)."""
    rng = np.random.default_rng(seed + 2003)
    d = np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0)
    out = []
    for k in range(n):
        ang = np.deg2rad(rng.uniform(-jitter_deg, jitter_deg, size=3)) if jitter_deg > 0 else np.zeros(3)
        R_wc = rot_xyz(*ang)
        out.append(se3_inv(se3_from_rt(R_wc, k * step * d)))
    return out
