"""Deterministic synthetic scenes and cameras for the benchmark configurations.

Random numbers come from SplitMix64 (one 64-bit counter stream per seed),
uniforms are the top 53 bits scaled to [0, 1), normals are Box-Muller
(cosine branch, two uniforms per normal) in float64, and every array is cast
to float32 at the end.  Draw order within a scene is fixed: means (3N
uniforms), log-scales (3N normals), quaternions (4N normals), opacity logits
(N normals), SH DC (3N normals), SH higher bands (3N(K-1) normals).

Layouts follow RustGS's DeviceSplats (training/engine/splats.rs:7-18):
transforms [N,10] = mean xyz, quaternion wxyz, log-scale xyz; sh [N,K,3]
coefficient-major with RGB interleaved; raw_opac [N] opacity logits.

Cameras follow RustGS's conventions: a world-to-camera matrix stored
column-major (core/camera.rs:30-32), x right / y down / z forward (the
projection of project_forward.wgsl:59 maps +y to increasing image rows),
principal point at the image centre (Intrinsics::from_focal,
rustscan-types/src/camera.rs:43-52).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Rng:
    """Counter-based SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    _CHUNK = 1 << 20

    def _mix(self, start: int, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            z = np.arange(start + 1, start + n + 1, dtype=np.uint64)
            z *= _GOLDEN
            z += self.seed
            z ^= z >> np.uint64(30)
            z *= _M1
            z ^= z >> np.uint64(27)
            z *= _M2
            z ^= z >> np.uint64(31)
        return z

    def u64(self, n: int) -> np.ndarray:
        out = np.empty(n, np.uint64)
        for a in range(0, n, self._CHUNK):
            b = min(n, a + self._CHUNK)
            out[a:b] = self._mix(self.counter + a, b - a)
        self.counter += n
        return out

    def uniform(self, n: int) -> np.ndarray:
        out = np.empty(n, np.float64)
        for a in range(0, n, self._CHUNK):
            b = min(n, a + self._CHUNK)
            z = self._mix(self.counter + a, b - a)
            z >>= np.uint64(11)
            out[a:b] = z
        out *= 1.0 / 9007199254740992.0
        self.counter += n
        return out

    def normal(self, n: int) -> np.ndarray:
        out = np.empty(n, np.float64)
        half = self._CHUNK // 2
        for a in range(0, n, half):
            b = min(n, a + half)
            u = self.uniform(2 * (b - a))
            u1, u2 = u[0::2], u[1::2]
            out[a:b] = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)
        return out


def sh_coeffs(degree: int) -> int:
    return (degree + 1) * (degree + 1)


@dataclass
class Splats:
    transforms: np.ndarray  # [N,10] float32
    sh: np.ndarray  # [N,K,3] float32
    raw_opac: np.ndarray  # [N] float32
    sh_degree: int

    @property
    def n(self) -> int:
        return int(self.raw_opac.shape[0])

    def copy(self) -> "Splats":
        return Splats(self.transforms.copy(), self.sh.copy(), self.raw_opac.copy(), self.sh_degree)


def gaussian_scene(
    n: int,
    seed: int,
    sh_degree: int,
    mean_lo: float = -1.0,
    mean_hi: float = 1.0,
    center=(0.0, 0.0, 0.0),
    log_scale_mu: float = -3.5,
    log_scale_sigma: float = 0.3,
    dc_sigma: float = 0.5,
    rest_sigma: float = 0.05,
    tail_fraction: float = 0.0,
    tail_seed: int | None = None,
    tail_factor: float = 20.0,
) -> Splats:
    """Means uniform in [lo,hi]^3 + center, log-scales N(mu, sigma) per axis,
    quaternions normalised N(0,1)^4 (wxyz), opacity logits N(0,1), SH DC
    N(0, dc_sigma) and higher bands N(0, rest_sigma).  With tail_fraction > 0 a
    seeded subset (stream tail_seed) gets log-scale + ln(tail_factor)."""
    rng = Rng(seed)
    k = sh_coeffs(sh_degree)
    means = mean_lo + (mean_hi - mean_lo) * rng.uniform(3 * n).reshape(n, 3) + np.asarray(center, np.float64)
    log_scales = log_scale_mu + log_scale_sigma * rng.normal(3 * n).reshape(n, 3)
    q = rng.normal(4 * n).reshape(n, 4)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    raw_opac = rng.normal(n)
    sh = np.zeros((n, k, 3), np.float64)
    sh[:, 0, :] = dc_sigma * rng.normal(3 * n).reshape(n, 3)
    if k > 1:
        sh[:, 1:, :] = rest_sigma * rng.normal(3 * n * (k - 1)).reshape(n, k - 1, 3)
    if tail_fraction > 0.0:
        trng = Rng(tail_seed if tail_seed is not None else seed + 1)
        sel = trng.uniform(n) < tail_fraction
        log_scales[sel] += math.log(tail_factor)
    transforms = np.concatenate([means, q, log_scales], axis=1).astype(np.float32)
    return Splats(transforms, sh.astype(np.float32), raw_opac.astype(np.float32), sh_degree)


@dataclass
class Camera:
    viewmat: np.ndarray  # [16] float32, column-major world->camera
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int
    cam_pos: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))

    def matrix(self) -> np.ndarray:
        return self.viewmat.reshape(4, 4).T.copy()


def pinhole_identity(width: int, height: int, focal: float) -> Camera:
    """Identity pose, Intrinsics::from_focal(f, W, H)."""
    vm = np.eye(4, dtype=np.float32).T.reshape(16).copy()
    return Camera(vm, np.float32(focal), np.float32(focal), np.float32(width / 2.0), np.float32(height / 2.0),
                  width, height, np.zeros(3, np.float32))


def look_at_camera(position, forward, width: int, height: int, focal: float) -> Camera:
    """Camera at `position` looking along `forward` with world +y up."""
    p = np.asarray(position, np.float64)
    f = np.asarray(forward, np.float64)
    f = f / np.linalg.norm(f)
    up = np.array([0.0, 1.0, 0.0])
    x = np.cross(f, up)
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    R = np.stack([x, y, f], axis=0)  # rows: camera axes in world coordinates
    t = -R @ p
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = t
    vm = M.T.reshape(16).astype(np.float32)  # column-major
    return Camera(vm, np.float32(focal), np.float32(focal), np.float32(width / 2.0), np.float32(height / 2.0),
                  width, height, p.astype(np.float32))


def ring_cameras(sizes, radius: float = 5.0, height: float = 1.0, jitter_deg: float = 10.0, seed: int = 7,
                 hfov_focal: float = 0.866) -> list[Camera]:
    """Cameras evenly spaced on a ring of `radius` at `height`, each looking at
    the origin with a uniform yaw jitter in [-jitter, +jitter] degrees;
    fx = fy = hfov_focal * W (0.866 W: 60 degree horizontal field of view)."""
    rng = Rng(seed)
    count = len(sizes)
    jit = (rng.uniform(count) * 2.0 - 1.0) * math.radians(jitter_deg)
    cams = []
    for i, (w, h) in enumerate(sizes):
        th = 2.0 * math.pi * i / count
        p = np.array([radius * math.sin(th), height, radius * math.cos(th)])
        f = -p / np.linalg.norm(p)
        c, s = math.cos(jit[i]), math.sin(jit[i])
        f = np.array([c * f[0] + s * f[2], f[1], -s * f[0] + c * f[2]])  # yaw about world y
        cams.append(look_at_camera(p, f, w, h, hfov_focal * w))
    return cams


# ---------------------------------------------------------------- configs
C4_SIZES = [(640, 480), (960, 540), (1280, 720), (1280, 720), (1280, 720), (1600, 900), (1920, 1080), (1920, 1080)]


def config1(n: int = 16384):
    """Plain forward render: N Gaussians, means U[-1,1]^3 + (0,0,4), SH0, 320x240, f=300."""
    sc = gaussian_scene(n, seed=1, sh_degree=0, center=(0.0, 0.0, 4.0))
    return sc, [pinhole_identity(320, 240, 300.0)]


def config2(n: int = 262144, width: int = 640, height: int = 480):
    """Single-view fwd+bwd: means U[-2,2]^3 + (0,0,6), SH3, 640x480, f=500; target U[0,1] (seed 3)."""
    sc = gaussian_scene(n, seed=2, sh_degree=3, mean_lo=-2.0, mean_hi=2.0, center=(0.0, 0.0, 6.0))
    cam = pinhole_identity(width, height, 500.0 * width / 640.0)
    target = Rng(3).uniform(width * height * 3).reshape(height, width, 3).astype(np.float32)
    return sc, [cam], [target]


def config3_scene(n: int = 1048576):
    """Training scene (init seed 5) and ground truth (seed 6): means U[-1,1]^3, SH3."""
    init = gaussian_scene(n, seed=5, sh_degree=3)
    gt = gaussian_scene(n, seed=6, sh_degree=3)
    return init, gt


def config3_cameras(count: int = 32, width: int = 1280, height: int = 720):
    return ring_cameras([(width, height)] * count, seed=7)


def config4_cameras():
    return ring_cameras(C4_SIZES, seed=7)


def config5_scene(n: int = 4194304):
    return gaussian_scene(n, seed=9, sh_degree=3, tail_fraction=0.01, tail_seed=8, tail_factor=20.0)


def config5_cameras(count: int = 16):
    return ring_cameras([(1920, 1080)] * count, seed=7)


# learning rates of the training configurations (constant schedule)
LR = dict(position=1.6e-4, scale=5e-3, rotation=1e-3, opacity=5e-2, sh_dc=2.5e-3, sh_rest=1.25e-4)
