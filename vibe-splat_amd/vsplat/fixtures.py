"""Seeded inputs: Gaussian clouds, cameras and RGBA8 targets.

Two families:

* The reference's own test fixtures (tests/reference_helpers.hpp:17-94):
  ``random_cloud``, ``train_camera``, ``smooth_target`` and ``grid_in_front``.
  They draw from ``std::mt19937`` through libstdc++'s
  ``uniform_real_distribution<float>`` (``generate_canonical<float, 24>``: one
  32-bit draw, ``float(u) / 2^32``, then ``a + (b - a) * c`` in float), which is
  reproduced here draw for draw. Only ``logf``/``sinf``/``cosf`` are taken from
  numpy in double and rounded, which can differ from glibc by one ulp.

* The benchmark workloads (the five ``CONFIGS``), with explicit RNG maths so
  that any implementation can regenerate them: SplitMix64 over a counter,
  53-bit uniforms, Box-Muller normals. ``draw(seed, stream, n)`` is the n-vector
  of uint64 outputs number ``stream * 2^32 + k`` of the generator seeded with
  ``seed``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SH_C0 = np.float32(0.28209479177387814)


@dataclass
class Cloud:
    """GaussianCloud (src/splat/gaussian_cloud.hpp:21-38): raw parameters."""

    positions: np.ndarray  # (N, 3) float32
    scales: np.ndarray  # (N, 3) log standard deviations
    rotations: np.ndarray  # (N, 4) quaternion w x y z, not normalised
    opacities: np.ndarray  # (N,) logit
    sh: np.ndarray  # (N, (d+1)^2, 3)
    sh_degree: int

    @property
    def count(self) -> int:
        return int(self.opacities.shape[0])

    @property
    def coefficients(self) -> int:
        return (self.sh_degree + 1) ** 2

    def copy(self) -> "Cloud":
        return Cloud(self.positions.copy(), self.scales.copy(), self.rotations.copy(), self.opacities.copy(),
                     self.sh.copy(), self.sh_degree)

    @staticmethod
    def zeros(count: int, degree: int) -> "Cloud":
        c = (degree + 1) ** 2
        return Cloud(np.zeros((count, 3), np.float32), np.zeros((count, 3), np.float32),
                     np.zeros((count, 4), np.float32), np.zeros(count, np.float32),
                     np.zeros((count, c, 3), np.float32), degree)


@dataclass
class Camera:
    """Camera (src/render/camera.hpp:9-18). view is world->camera, a 4x4 whose
    column-major flattening is what the C interfaces take."""

    view: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))
    position: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    fx: float = 1.0
    fy: float = 1.0
    cx: float = 0.0
    cy: float = 0.0
    width: int = 1
    height: int = 1

    def view_column_major(self) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(self.view, np.float32).T.reshape(16))


# ---------------------------------------------------------------- reference fixtures

class StdMt19937:
    """std::mt19937(seed): numpy's legacy RandomState uses the same
    initialisation (init_genrand) and, for a full-range uint32 request, hands
    out the raw 32-bit outputs in order."""

    def __init__(self, seed: int):
        self._state = np.random.RandomState(seed & 0xFFFFFFFF)

    def raw(self, n: int) -> np.ndarray:
        return self._state.randint(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32) if n > 0 else np.zeros(0, np.uint32)

    def canonical_float(self, n: int) -> np.ndarray:
        """generate_canonical<float, 24>: float(u) / 2^32, kept below 1."""
        c = self.raw(n).astype(np.float32) / np.float32(4294967296.0)
        c[c >= np.float32(1.0)] = np.nextafter(np.float32(1.0), np.float32(0.0))
        return c


def _uniform_float(c: np.ndarray, a: float, b: float) -> np.ndarray:
    return (c * np.float32(np.float32(b) - np.float32(a))) + np.float32(a)


def _logf(x: np.ndarray) -> np.ndarray:
    return np.log(x.astype(np.float64)).astype(np.float32)


def random_cloud(count: int, degree: int, seed: int) -> Cloud:
    """tests/reference_helpers.hpp:17-34, draw for draw."""
    rng = StdMt19937(seed)
    coefficients = (degree + 1) ** 2
    per = 3 + 3 + 4 + 1 + 3 + (coefficients * 3 - 3)
    c = rng.canonical_float(count * per).reshape(count, per)
    cloud = Cloud.zeros(count, degree)
    cloud.positions[:, 0] = _uniform_float(c[:, 0], -1, 1) * np.float32(1.4)
    cloud.positions[:, 1] = _uniform_float(c[:, 1], -1, 1) * np.float32(1.1)
    cloud.positions[:, 2] = np.float32(3.0) + c[:, 2] * np.float32(3.0)
    cloud.scales[:] = _logf(np.float32(0.06) + np.float32(0.2) * c[:, 3:6])
    cloud.rotations[:] = _uniform_float(c[:, 6:10], -1, 1)
    cloud.opacities[:] = _uniform_float(c[:, 10], -1, 1) * np.float32(1.5)
    sh = cloud.sh.reshape(count, coefficients * 3)
    sh[:, 0:3] = (c[:, 11:14] - np.float32(0.5)) / SH_C0
    if coefficients > 1:
        sh[:, 3:] = _uniform_float(c[:, 14:], -1, 1) * np.float32(0.2)
    return cloud


def train_camera(width: int = 64, height: int = 48) -> Camera:
    """tests/reference_helpers.hpp:37-45: at the origin, looking along +z."""
    return Camera(fx=60.0, fy=60.0, cx=float(np.float32(width) * np.float32(0.5)),
                  cy=float(np.float32(height) * np.float32(0.5)), width=width, height=height)


def smooth_target(width: int, height: int, seed: int) -> np.ndarray:
    """tests/reference_helpers.hpp:48-61: RGBA8, (H, W, 4)."""
    rng = StdMt19937(seed)
    draws = rng.raw(width * height * 2).reshape(height, width, 2) % 16
    x = np.arange(width, dtype=np.float32)[None, :]
    y = np.arange(height, dtype=np.float32)[:, None]
    sx = np.sin((np.float32(0.2) * x).astype(np.float64)).astype(np.float32)
    cy = np.cos((np.float32(0.15) * y).astype(np.float64)).astype(np.float32)
    out = np.zeros((height, width, 4), np.uint8)
    out[..., 0] = (np.float32(127.0) + np.float32(100.0) * sx + draws[..., 0].astype(np.float32)).astype(np.uint8)
    out[..., 1] = (np.float32(127.0) + np.float32(100.0) * cy + draws[..., 1].astype(np.float32)).astype(np.uint8)
    xi = np.arange(width)[None, :]
    yi = np.arange(height)[:, None]
    out[..., 2] = (60 + (xi + yi) % 120).astype(np.uint8)
    out[..., 3] = 255
    return out


def grid_in_front(camera: Camera, side: int) -> Cloud:
    """tests/reference_helpers.hpp:78-94: a side x side grid of grey Gaussians 2 units ahead."""
    cloud = Cloud.zeros(side * side, 0)
    to_world = np.asarray(camera.view, np.float32)[:3, :3].T
    spacing = np.float32(2.6) / np.float32(side)
    i = np.arange(side * side)
    x = ((i % side).astype(np.float32) + np.float32(0.5) - np.float32(side) * np.float32(0.5)) * spacing
    y = ((i // side).astype(np.float32) + np.float32(0.5) - np.float32(side) * np.float32(0.5)) * spacing
    local = np.stack([x, y, np.full_like(x, 2.0)], 1)
    cloud.positions[:] = camera.position[None, :] + local @ to_world.T
    cloud.scales[:] = _logf(np.full((side * side, 3), spacing * np.float32(0.75), np.float32))
    cloud.rotations[:, 0] = 1.0
    return cloud


# ---------------------------------------------------------------- explicit RNG maths

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def draw(seed: int, stream: int, n: int) -> np.ndarray:
    """SplitMix64: output k of the generator whose state starts at
    seed * GAMMA + (stream * 2^32) * GAMMA; state advances by GAMMA per output."""
    with np.errstate(over="ignore"):
        start = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * _GAMMA + np.uint64((stream << 32) & 0xFFFFFFFFFFFFFFFF) * _GAMMA
        z = start + (np.arange(1, n + 1, dtype=np.uint64) * _GAMMA)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """53-bit uniforms in [0, 1), as float64."""
    return (draw(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Box-Muller on pairs of uniforms: u1 in (0, 1] from 1 - uniform."""
    m = (n + 1) // 2
    u = uniform(seed, stream, 2 * m).reshape(m, 2)
    r = np.sqrt(-2.0 * np.log(1.0 - u[:, 0]))
    t = 2.0 * math.pi * u[:, 1]
    return np.stack([r * np.cos(t), r * np.sin(t)], 1).reshape(-1)[:n]


def look_at(eye, target, up, vfov_deg: float, width: int, height: int) -> Camera:
    """Camera::look_at (src/render/camera.cpp:8-37), in float."""
    f32 = np.float32
    eye = np.asarray(eye, f32)
    target = np.asarray(target, f32)
    up = np.asarray(up, f32)

    def normalize(v):
        return (v * f32(1.0 / np.sqrt(np.float64(np.dot(v, v))))).astype(f32)

    forward = normalize(target - eye)
    right = np.cross(-up, forward).astype(f32)
    if float(np.dot(right, right)) < 1e-12:
        right = np.array([1, 0, 0], f32)
    right = normalize(right)
    down = np.cross(forward, right).astype(f32)
    view = np.eye(4, dtype=f32)
    view[0, :3] = right
    view[1, :3] = down
    view[2, :3] = forward
    view[0, 3] = -np.dot(right, eye)
    view[1, 3] = -np.dot(down, eye)
    view[2, 3] = -np.dot(forward, eye)
    fy = f32(0.5) * f32(height) / f32(math.tan(math.radians(vfov_deg) * 0.5))
    return Camera(view=view, position=eye.copy(), fx=float(fy), fy=float(fy), cx=float(f32(0.5) * f32(width)),
                  cy=float(f32(0.5) * f32(height)), width=width, height=height)


def synthetic_cloud(count: int, degree: int, seed: int, *, log_scale=(-4.0, 0.5), opacity_logit=(0.0, 1.0),
                    sh_sigma: float | None = None, blobs: int = 0, blob_sigma: float = 0.2) -> Cloud:
    """The benchmark clouds. Streams: 0 means, 1 log-scales, 2 quaternions,
    3 opacity logits, 4 colours / SH, 5 blob centres, 6 blob choice."""
    cloud = Cloud.zeros(count, degree)
    if blobs > 0:
        centres = uniform(seed, 5, blobs * 3).reshape(blobs, 3) * 2.0 - 1.0
        choice = np.minimum((uniform(seed, 6, count) * blobs).astype(np.int64), blobs - 1)
        cloud.positions[:] = centres[choice] + blob_sigma * normal(seed, 0, count * 3).reshape(count, 3)
    else:
        cloud.positions[:] = uniform(seed, 0, count * 3).reshape(count, 3) * 2.0 - 1.0
    cloud.scales[:] = log_scale[0] + log_scale[1] * normal(seed, 1, count * 3).reshape(count, 3)
    q = normal(seed, 2, count * 4).reshape(count, 4)
    cloud.rotations[:] = q / np.linalg.norm(q, axis=1, keepdims=True)
    cloud.opacities[:] = opacity_logit[0] + opacity_logit[1] * normal(seed, 3, count)
    c = cloud.coefficients
    if sh_sigma is None:
        rgb = uniform(seed, 4, count * 3).reshape(count, 3)
        cloud.sh[:, 0, :] = (rgb - 0.5) / float(SH_C0)
    else:
        cloud.sh[:] = sh_sigma * normal(seed, 4, count * c * 3).reshape(count, c, 3)
    return cloud


def random_target(width: int, height: int, seed: int) -> np.ndarray:
    """Uniform random RGBA8, alpha 255."""
    out = (draw(seed, 7, width * height * 4) >> np.uint64(56)).astype(np.uint8).reshape(height, width, 4)
    out[..., 3] = 255
    return out


def smooth_target_explicit(width: int, height: int, seed: int) -> np.ndarray:
    """The smooth_target formula (tests/reference_helpers.hpp:48-61) with the explicit RNG."""
    d = (draw(seed, 8, width * height * 2) % np.uint64(16)).astype(np.float32).reshape(height, width, 2)
    x = np.arange(width, dtype=np.float32)[None, :]
    y = np.arange(height, dtype=np.float32)[:, None]
    out = np.zeros((height, width, 4), np.uint8)
    out[..., 0] = (127.0 + 100.0 * np.sin(0.2 * x) + d[..., 0]).astype(np.uint8)
    out[..., 1] = (127.0 + 100.0 * np.cos(0.15 * y) + d[..., 1]).astype(np.uint8)
    out[..., 2] = (60 + (np.arange(width)[None, :] + np.arange(height)[:, None]) % 120).astype(np.uint8)
    out[..., 3] = 255
    return out


@dataclass
class Workload:
    name: str
    cloud: Cloud
    cameras: list
    targets: list
    lambda_ssim: float
    fixed_point: bool
    description: str


RING_SIZES = [(640, 480), (960, 540), (1280, 720), (1280, 720), (1280, 720), (1600, 900), (1920, 1080), (1920, 1080)]


def config(index: int, *, count: int | None = None) -> Workload:
    """The five workloads, seeded 1000 + index. `count` overrides N (tests)."""
    seed = 1000 + index
    up = (0.0, -1.0, 0.0)
    if index == 0:
        n = count or 10_000
        cloud = synthetic_cloud(n, 0, seed)
        cam = look_at((0, 0, -3), (0, 0, 0), up, 60.0, 256, 256)
        return Workload("c0: 10k SH0 256x256 L1", cloud, [cam], [random_target(256, 256, seed)], 0.0, False,
                        "10k Gaussians, SH degree 0, 256x256, pure L1 against a uniform random target")
    if index == 1:
        n = count or 1_000_000
        cloud = synthetic_cloud(n, 3, seed, sh_sigma=0.3)
        cam = look_at((0, 0, -3), (0, 0, 0), up, 60.0, 1280, 720)
        return Workload("c1: 1M SH3 1280x720 L1+SSIM fixed-point", cloud, [cam],
                        [smooth_target_explicit(1280, 720, seed)], 0.2, True,
                        "1M Gaussians, SH degree 3 (48 coefficients N(0,0.3)), 1280x720, 0.8 L1 + 0.2 D-SSIM, fixed-point backward")
    if index == 2:
        n = count or 3_000_000
        cloud = synthetic_cloud(n, 3, seed, sh_sigma=0.3, blobs=64)
        cam = look_at((0, 0, -3), (0, 0, 0), up, 60.0, 1920, 1080)
        return Workload("c2: 3M SH3 64 blobs 1920x1080", cloud, [cam], [smooth_target_explicit(1920, 1080, seed)], 0.2,
                        True, "3M Gaussians in 64 blobs, SH degree 3, 1920x1080")
    if index == 3:
        n = count or 1_000_000
        cloud = synthetic_cloud(n, 3, seed, sh_sigma=0.3)
        cams, targets = [], []
        for k, (w, h) in enumerate(RING_SIZES):
            th = 2.0 * math.pi * k / 8.0
            cams.append(look_at((3.0 * math.sin(th), 0.0, -3.0 * math.cos(th)), (0, 0, 0), up, 60.0, w, h))
            targets.append(smooth_target_explicit(w, h, seed * 16 + k))
        return Workload("c3: 1M SH3 8 views ring", cloud, cams, targets, 0.2, True,
                        "1M Gaussians, SH degree 3, 8 views of uneven size on a ring of radius 3")
    if index == 4:
        n = count or 6_000_000
        cloud = synthetic_cloud(n, 3, seed, sh_sigma=0.3, log_scale=(-3.5, 1.2), opacity_logit=(1.0, 1.5))
        cam = look_at((0, 0, -3), (0, 0, 0), up, 60.0, 1920, 1080)
        return Workload("c4: 6M heavy tails 1920x1080", cloud, [cam], [smooth_target_explicit(1920, 1080, seed)], 0.2,
                        True, "6M Gaussians, log-scales N(-3.5,1.2), opacity logit N(1,1.5), 1920x1080")
    raise ValueError(f"no config {index}")


# ---------------------------------------------------------------- SH packing

def float_to_half_bits(x: np.ndarray) -> np.ndarray:
    """IEEE binary16 bits, round to nearest even (numpy's float16 cast)."""
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def pack_half_sh(sh: np.ndarray) -> np.ndarray:
    """pack_half_sh (src/splat/gaussian_cloud.cpp:537-550): per Gaussian the DC
    term's 3 floats, then the other coefficients as halves, two to a word, low half first."""
    n = sh.shape[0]
    flat = np.asarray(sh, np.float32).reshape(n, -1)
    floats = flat.shape[1]
    words = 3 + (floats - 3 + 1) // 2
    out = np.zeros((n, words), np.uint32)
    out[:, :3] = flat[:, :3].view(np.uint32)
    rest = flat[:, 3:]
    if rest.shape[1] % 2:
        rest = np.concatenate([rest, np.zeros((n, 1), np.float32)], 1)
    halves = float_to_half_bits(rest).astype(np.uint32)
    if floats - 3 < rest.shape[1]:
        halves[:, -1] = 0
    out[:, 3:] = halves[:, 0::2] | (halves[:, 1::2] << np.uint32(16))
    return out


def unpack_half_sh(words: np.ndarray, coefficients: int) -> np.ndarray:
    n = words.shape[0]
    floats = 3 * coefficients
    out = np.zeros((n, floats), np.float32)
    out[:, :3] = words[:, :3].view(np.float32)
    if floats > 3:
        lo = (words[:, 3:] & np.uint32(0xFFFF)).astype(np.uint16).view(np.float16).astype(np.float32)
        hi = (words[:, 3:] >> np.uint32(16)).astype(np.uint16).view(np.float16).astype(np.float32)
        inter = np.stack([lo, hi], 2).reshape(n, -1)
        out[:, 3:] = inter[:, : floats - 3]
    return out.reshape(n, coefficients, 3)


def half_sh_words(coefficients: int) -> int:
    """half_sh_words (src/splat/gaussian_cloud.hpp:44-46)."""
    return 3 + (3 * (coefficients - 1) + 1) // 2


def bf16_sh_words(coefficients: int) -> int:
    """Words per Gaussian when every SH coefficient is a bf16, two to a word."""
    return (3 * coefficients + 1) // 2


def float_to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """bf16 bits, round to nearest even on the float32 bits (the trainer's upload rule)."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint32) & np.uint32(0xFFFF)


def pack_bf16_sh(sh: np.ndarray) -> np.ndarray:
    """Every coefficient as bf16, two to a word, low half first."""
    n = sh.shape[0]
    flat = np.asarray(sh, np.float32).reshape(n, -1)
    if flat.shape[1] % 2:
        flat = np.concatenate([flat, np.zeros((n, 1), np.float32)], 1)
    bits = float_to_bf16_bits(flat)
    if flat.shape[1] > np.asarray(sh).reshape(n, -1).shape[1]:
        bits[:, -1] = 0
    return (bits[:, 0::2] | (bits[:, 1::2] << np.uint32(16))).astype(np.uint32)


def unpack_bf16_sh(words: np.ndarray, coefficients: int) -> np.ndarray:
    n = words.shape[0]
    lo = (words << np.uint32(16)).view(np.float32)
    hi = (words & np.uint32(0xFFFF0000)).view(np.float32)
    inter = np.stack([lo, hi], 2).reshape(n, -1)
    return np.ascontiguousarray(inter[:, : 3 * coefficients]).reshape(n, coefficients, 3)


def sh_words(coefficients: int, storage: int = 0) -> int:
    return bf16_sh_words(coefficients) if storage == 1 else half_sh_words(coefficients)


def pack_sh(sh: np.ndarray, storage: int = 0) -> np.ndarray:
    """The trainer's SH words: storage 0 the reference's half layout, 1 all bf16."""
    return pack_bf16_sh(sh) if storage == 1 else pack_half_sh(sh)


def unpack_sh(words: np.ndarray, coefficients: int, storage: int = 0) -> np.ndarray:
    return unpack_bf16_sh(words, coefficients) if storage == 1 else unpack_half_sh(words, coefficients)
