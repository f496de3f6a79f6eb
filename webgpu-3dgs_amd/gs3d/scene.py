"""Synthetic Gaussian-splat scenes and the loader's activations.

The generator produces *pre-activation* values, i.e. what a 3DGS PLY row holds
(x, y, z, scale_0..2 as log-scales, rot_0..3 raw, opacity as a logit, f_dc_0..2,
f_rest_k), rounded to float32 as a PLY stores them. :func:`activate` then applies
the reference loader's activations (``src/workers/SceneWorker.ts:407-464``) so
the same code path serves PLY files and synthetic scenes.

Randomness is counter-based (SplitMix64 over ``base(seed, stream) + i``), so a
scene is a pure function of (config, index) and does not depend on numpy's
generator versions. Normals come from Box-Muller on two uniform streams.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SH_C0 = 0.28209479177387814  # SceneWorker.ts:428

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_MASK64 = (1 << 64) - 1


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _mix64_int(x: int) -> int:
    x &= _MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _MASK64
    return x ^ (x >> 31)


def uniform(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """n float64 uniforms in [0, 1): SplitMix64 outputs of state base + (i+1)*golden."""
    base = _mix64_int(seed * 0x100000001B3 + stream * 0x9E3779B1 + 0x632BE59BD9B4E019)
    i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = _mix64(np.uint64(base) + i * _GOLDEN)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int, mean: float = 0.0, std: float = 1.0) -> np.ndarray:
    """n float64 normals (Box-Muller on streams 2*stream and 2*stream+1)."""
    u1 = uniform(seed, 2 * stream + 1000, n)
    u2 = uniform(seed, 2 * stream + 1001, n)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return mean + std * r * np.cos(2.0 * math.pi * u2)


@dataclass
class RawSplats:
    """Pre-activation PLY-equivalent splat attributes (float32)."""

    xyz: np.ndarray  # (N, 3)
    log_scale: np.ndarray  # (N, 3)
    rot: np.ndarray  # (N, 4) raw quaternion (w, x, y, z)
    opacity_logit: np.ndarray  # (N,)
    f_dc: np.ndarray  # (N, 3)
    f_rest: np.ndarray | None = None  # (N, 3*R) in f_rest_k index order

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def sh_degree(self) -> int:
        """SceneWorker.inferSphericalHarmonicsDegree (SceneWorker.ts:501-506)."""
        f = 0 if self.f_rest is None else self.f_rest.shape[1]
        return 3 if f >= 45 else 2 if f >= 24 else 1 if f >= 9 else 0


@dataclass
class Splats:
    """Activated splats in the loader's vec4 layouts (GaussianSplatData, types.ts:14-22)."""

    positions: np.ndarray  # (N, 4) f32 (x, y, z, 1)
    colors: np.ndarray  # (N, 4) f32 (clamp01(0.5 + C0 f_dc), clamp01(sigmoid(opacity)))
    scales: np.ndarray  # (N, 4) f32 (e^s0, e^s1, e^s2, 0)
    rotations: np.ndarray  # (N, 4) f32 normalised (w, x, y, z)
    f_dc: np.ndarray  # (N, 3) f32 raw f_dc (kept for SH evaluation)
    sh_rest: np.ndarray | None  # (N, 3*R) f32 or None
    sh_degree: int = 0
    bounds: tuple = field(default_factory=tuple)

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])


def activate(raw: RawSplats) -> Splats:
    """The loader's per-row activations, vectorised (SceneWorker.ts:407-464).

    Arithmetic is float64 from float32 inputs (JS numbers), stored as float32.
    """
    n = raw.n
    xyz = raw.xyz.astype(np.float64)
    pos = np.empty((n, 4), np.float32)
    pos[:, :3] = xyz
    pos[:, 3] = 1.0
    col = np.empty((n, 4), np.float32)
    col[:, :3] = np.clip(0.5 + SH_C0 * raw.f_dc.astype(np.float64), 0.0, 1.0)
    op = raw.opacity_logit.astype(np.float64)
    col[:, 3] = np.clip(1.0 / (1.0 + np.exp(-op)), 0.0, 1.0)
    sc = np.empty((n, 4), np.float32)
    sc[:, :3] = np.exp(raw.log_scale.astype(np.float64))
    sc[:, 3] = 0.0
    r = raw.rot.astype(np.float64)
    length = np.sqrt(np.sum(r * r, axis=1))
    inv = np.where(length > 0.0, 1.0 / np.where(length > 0.0, length, 1.0), 1.0)
    rot = (r * inv[:, None]).astype(np.float32)
    deg = raw.sh_degree
    rest = None
    if deg > 0:
        ncoef = 3 * ((deg + 1) ** 2 - 1)
        rest = np.ascontiguousarray(raw.f_rest[:, :ncoef], dtype=np.float32)
    bounds = ()
    if n:
        bounds = (tuple(xyz.min(axis=0)), tuple(xyz.max(axis=0)))
    return Splats(pos, col, sc, rot, np.ascontiguousarray(raw.f_dc, np.float32), rest, deg, bounds)


def _f32(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def uniform_cube(seed: int, n: int, log_scale_mean: float = -4.5, log_scale_std: float = 0.5,
                 sh_degree: int = 0) -> RawSplats:
    """Positions uniform in [-1,1]^3 and the attribute law of the 100k SH0 workload."""
    xyz = np.stack([uniform(seed, s, n) * 2.0 - 1.0 for s in range(3)], axis=1)
    return _attributes(seed, xyz, log_scale_mean, log_scale_std, sh_degree)


def clustered(seed: int, n: int, clusters: int, log_scale_mean: float = -4.5,
              log_scale_std: float = 0.5, sh_degree: int = 3) -> RawSplats:
    """Mixture of `clusters` anisotropic Gaussians: centres uniform in [-4,4]^3,
    per-axis std U(0.1, 1.0); attributes as :func:`uniform_cube`, f_rest N(0, 0.1)."""
    centres = np.stack([uniform(seed, 10 + a, clusters) * 8.0 - 4.0 for a in range(3)], axis=1)
    stds = np.stack([0.1 + 0.9 * uniform(seed, 13 + a, clusters) for a in range(3)], axis=1)
    cid = np.minimum((uniform(seed, 16, n) * clusters).astype(np.int64), clusters - 1)
    z = np.stack([normal(seed, 20 + a, n) for a in range(3)], axis=1)
    xyz = centres[cid] + stds[cid] * z
    return _attributes(seed, xyz, log_scale_mean, log_scale_std, sh_degree)


def _attributes(seed, xyz, ls_mean, ls_std, sh_degree) -> RawSplats:
    n = xyz.shape[0]
    log_scale = np.stack([normal(seed, 30 + a, n, ls_mean, ls_std) for a in range(3)], axis=1)
    rot = np.stack([normal(seed, 40 + a, n) for a in range(4)], axis=1)
    opac = normal(seed, 50, n, 0.0, 1.5)
    f_dc = np.stack([normal(seed, 60 + a, n, 0.0, 0.5) for a in range(3)], axis=1)
    f_rest = None
    if sh_degree > 0:
        r = 3 * ((sh_degree + 1) ** 2 - 1)
        f_rest = np.stack([normal(seed, 100 + k, n, 0.0, 0.1) for k in range(r)], axis=1)
        f_rest = _f32(f_rest)
    return RawSplats(_f32(xyz), _f32(log_scale), _f32(rot), _f32(opac), _f32(f_dc), f_rest)


@dataclass
class Workload:
    """A named render workload: splats, camera placements and resolutions."""

    name: str
    raw: RawSplats
    views: list  # list of (eye xyz, width, height)
    tile: int = 16
    fov_y: float = math.pi / 3.0
    orbit_radius: float = 0.0  # the orbit the single view sits on (0: not an orbit camera)


def orbit_eye(radius: float, azimuth_deg: float, elevation_deg: float) -> tuple:
    az, el = math.radians(azimuth_deg), math.radians(elevation_deg)
    return (radius * math.cos(el) * math.sin(az), radius * math.sin(el),
            radius * math.cos(el) * math.cos(az))


RESOLUTIONS = ((640, 360), (1280, 720), (1920, 1080))


def workload(index: int, n: int | None = None, views: int | None = None) -> Workload:
    """The five workloads by index; `n` / `views` shrink them for tests.

    0: 100k uniform cube SH0, camera (0,0,3), 1280x720.
    1: 1M splats, 64 clusters, SH3, 1920x1080, orbit radius 8 (azimuth = elevation = 0).
    2: 6M splats, 256 clusters, log-scale N(-4.0, 0.7), SH3, 3840x2160, orbit radius 8.
    3: 2M splats, 128 clusters, SH3, 64 seeded orbit views of uneven resolution.
    4: 3M splats, 192 clusters, SH3, 32 orbit views at 1920x1080 (learned colour).
    """
    if index == 0:
        raw = uniform_cube(0, n or 100_000)
        return Workload("uniform-100k-sh0-720p", raw, [((0.0, 0.0, 3.0), 1280, 720)])
    if index == 1:
        raw = clustered(1, n or 1_000_000, 64)
        return Workload("clusters-1M-sh3-1080p", raw, [(orbit_eye(8.0, 0.0, 0.0), 1920, 1080)],
                        orbit_radius=8.0)
    if index == 2:
        raw = clustered(2, n or 6_000_000, 256, -4.0, 0.7)
        return Workload("clusters-6M-sh3-2160p", raw, [(orbit_eye(8.0, 0.0, 0.0), 3840, 2160)],
                        orbit_radius=8.0)
    if index == 3:
        raw = clustered(3, n or 2_000_000, 128)
        return Workload("clusters-2M-sh3-64views", raw, orbit_views(3, views or 64))
    if index == 4:
        raw = clustered(4, n or 3_000_000, 192)
        vs = orbit_views(4, views or 32, fixed=(1920, 1080))
        return Workload("clusters-3M-sh3-32views-learned", raw, vs)
    raise ValueError(f"no workload {index}")


def orbit_views(seed: int, count: int, fixed: tuple | None = None) -> list:
    """Seeded orbit: radius U(5,10), elevation U(-30,60) deg, azimuth U(0,360) deg,
    resolution uniform over RESOLUTIONS unless `fixed`."""
    r = 5.0 + 5.0 * uniform(seed, 900, count)
    el = -30.0 + 90.0 * uniform(seed, 901, count)
    az = 360.0 * uniform(seed, 902, count)
    ri = np.minimum((uniform(seed, 903, count) * 3).astype(np.int64), 2)
    out = []
    for i in range(count):
        w, h = fixed if fixed else RESOLUTIONS[int(ri[i])]
        out.append((orbit_eye(float(r[i]), float(az[i]), float(el[i])), w, h))
    return out
