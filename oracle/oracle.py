"""Test infrastructure: ctypes access to the two CPU oracles (liboracle.so).

Only the tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
this module; the product never does. Build with ``make -C oracle``.

Oracle A (``render``, ``loss``, ``gradients``, ``contributions``) is the
double-precision restatement of the reference's src/train/reference.cpp.
Oracle B (``Model``) is the single-precision model of the GPU pipeline.
``gather`` restates the state gather of a changed cloud (train_gather.comp);
``exposed_image``, ``exposure_backward``, ``score_fields``, ``error_mask`` and
``score_statistics`` restate the exposure and scoring passes elementwise in
float32, in this build's operation order; ``marked_pixels`` (Oracle A) is
reference.cpp's FP64 count of high-error pixels per Gaussian. The float32
restatements are pinned by tests/test_scoring.py (closed forms: a neutral
exposure is the identity, the exposure gradient against FP64 central
differences of Oracle A's loss, the edge weight of flat and stepped targets,
the error mask's threshold) and, on the GPU, by Oracle A's marked pixels;
it is pinned by tests/test_densify.py: a brute-force loop on small tables, the
identity table, the composition of two tables, and fresh (zero) state for
negative sources. ``position_noise`` restates train_noise.comp in float64
from the shader's float32 uniforms; tests/test_noise.py pins it by the
statistics the reference's own test checks (tests/test_train.cpp:322-350),
the normals' moments and the covariance of the moves (Sigma^2 for a rotated,
anisotropic Gaussian), and the gate's closed form.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class CCloud(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("scales", C.c_void_p), ("rotations", C.c_void_p), ("opacities", C.c_void_p),
                ("sh", C.c_void_p), ("count", C.c_uint32), ("sh_degree", C.c_int32)]


class CCamera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("position", C.c_float * 3), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("width", C.c_int32), ("height", C.c_int32)]


class COptions(C.Structure):
    _fields_ = [("background", C.c_float * 3), ("near_plane", C.c_float), ("no_cutoffs", C.c_int32),
                ("opacity_reg", C.c_float), ("scale_reg", C.c_float), ("mip_filter_2d", C.c_int32),
                ("filter3d", C.c_void_p), ("filter3d_count", C.c_uint32)]


class OBCamera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("position", C.c_float * 3), ("near_plane", C.c_float), ("fx", C.c_float),
                ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float), ("tan_half_x", C.c_float),
                ("tan_half_y", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32), ("count", C.c_uint32),
                ("sh_degree", C.c_uint32), ("sh_words", C.c_uint32), ("sh_coefficients", C.c_uint32),
                ("sh_format", C.c_uint32)]


class OBPreOptions(C.Structure):
    _fields_ = [("mip_filter_2d", C.c_int32), ("sh_dropout", C.c_float), ("sh_retained", C.c_uint32),
                ("sh_seed", C.c_uint32)]


class OBChainOptions(C.Structure):
    _fields_ = [("opacity_reg", C.c_float), ("scale_reg", C.c_float), ("mip_filter_2d", C.c_int32),
                ("sh_dropout", C.c_float), ("sh_retained", C.c_uint32), ("sh_seed", C.c_uint32),
                ("mode", C.c_uint32), ("flags", C.c_uint32), ("rates", C.c_float * 5), ("rest_scale", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float), ("correction1", C.c_float),
                ("correction2", C.c_float), ("seed", C.c_uint32), ("sh_floats", C.c_uint32), ("trust_region", C.c_float),
                ("curvature_scale", C.c_float), ("curvature_beta", C.c_float), ("per_value", C.c_float),
                ("pose_terms", C.c_void_p)]


CURVATURE_BETA = np.float32(0.999)
# Fixed-point steps per unit of a wave total (2^20), and the curvature pass's
# scale: its adjoint is +-1 per value, without the loss's 1 / (3 W H).
FIXED_STEPS = np.float32(1048576.0)
CURVATURE_FIXED_SCALE = float(FIXED_STEPS)


def curvature_scale(updates: int) -> float:
    """1 / (1 - beta^k) for the curvature's running average after k updates
    (trainer.cpp:460-461), beta^k kept in float32 by repeated multiplication."""
    b = np.float32(1.0)
    for _ in range(max(updates, 1)):
        b = np.float32(b * CURVATURE_BETA)
    return float(np.float32(1.0) / np.float32(np.float32(1.0) - b))


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise ImportError(f"{path} is not built: run make -C oracle")
        _LIB = C.CDLL(str(path))
        L = _LIB
        vp = C.c_void_p
        L.oa_render.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp]
        L.oa_contributions.argtypes = [vp, vp, vp, C.c_int, vp]
        L.oa_marked_pixels.argtypes = [vp, vp, vp, C.c_double, vp, C.c_int, vp]
        L.oa_loss.argtypes = [vp, vp, C.c_int, C.c_int, C.c_double, C.c_int, vp, vp]
        L.oa_gradients.argtypes = [vp, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.ob_preprocess.argtypes = [vp] * 11
        L.ob_stable_sort.argtypes = [vp, vp, C.c_uint32]
        L.ob_bin.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, vp, C.c_uint64, vp]
        L.ob_tile_ranges.argtypes = [vp, C.c_uint64, C.c_uint32, vp]
        L.ob_forward.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp]
        L.ob_loss.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_float, vp, vp, vp]
        L.ob_backward_fixed.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp, vp, C.c_float, C.c_uint32, vp, vp]
        L.ob_backward_float.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp, vp, C.c_uint32, vp, vp]
        L.ob_fixed_to_float.argtypes = [vp, C.c_uint64, C.c_float, vp]
        L.ob_chain.argtypes = [vp] * 21
        L.ob_sh_step.argtypes = [vp] * 5
        L.ob_adam.argtypes = [vp] * 15 + [C.c_uint32, C.c_uint32, C.c_uint32, vp]
        L.ob_exp_array.argtypes = [vp, vp, C.c_uint64]
        L.ob_log_array.argtypes = [vp, vp, C.c_uint64]
        L.ob_exp_array.restype = None
        L.ob_log_array.restype = None
        L.ob_hash_array.argtypes = [vp, C.c_uint32, C.c_uint64, vp]
        L.ob_uniform_array.argtypes = [vp, C.c_uint32, C.c_uint64, vp]
        L.ob_bf16_stochastic_array.argtypes = [vp, vp, C.c_uint32, C.c_uint64, vp]
        L.ob_half_stochastic_array.argtypes = [vp, vp, C.c_uint32, C.c_uint64, vp]
        for f in (L.ob_hash_array, L.ob_uniform_array, L.ob_bf16_stochastic_array, L.ob_half_stochastic_array):
            f.restype = None
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


NEUTRAL_EXPOSURE = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)


def exposed_image(image: np.ndarray, transform) -> np.ndarray:
    """The image as the view's exposure makes it (train_exposure.comp:17-24):
    rgb' = T (rgb, 1), each row's dot product added left to right in float32;
    the fourth value copied."""
    t = np.asarray(transform, np.float32).reshape(3, 4)
    img = np.asarray(image, np.float32).reshape(-1, 4)
    out = img.copy()
    for row in range(3):
        a = t[row]
        out[:, row] = ((a[0] * img[:, 0] + a[1] * img[:, 1]) + a[2] * img[:, 2]) + a[3]
    return out


def _tile_sums(values: np.ndarray) -> np.ndarray:
    """Per 256-value group, the sum in the GPU's workgroup tree: within each
    64-lane wave lane l plus lane l ^ m for m = 32 ... 1, then the four waves
    added in order."""
    n = values.shape[0]
    groups = (n + 255) // 256
    v = np.zeros((groups * 256,) + values.shape[1:], np.float32)
    v[:n] = values
    v = v.reshape((groups, 4, 64) + values.shape[1:])
    lanes = np.arange(64)
    for m in (32, 16, 8, 4, 2, 1):
        v = (v + v[:, :, lanes ^ m]).astype(np.float32)
    w = v[:, :, 0]
    return (((w[:, 0] + w[:, 1]).astype(np.float32) + w[:, 2]).astype(np.float32) + w[:, 3]).astype(np.float32)


def exposure_backward(image: np.ndarray, adjoint: np.ndarray, transform):
    """The adjoint taken back through the exposure and the twelve sums of
    dL/dT per group of 256 pixels (train_exposure_backward.comp:22-40):
    back = A^T g added row by row; totals g[row] (rgb, 1)[column]."""
    t = np.asarray(transform, np.float32).reshape(3, 4)
    img = np.asarray(image, np.float32).reshape(-1, 4)
    adj = np.asarray(adjoint, np.float32).reshape(-1, 4).copy()
    g = adj[:, :3].copy()
    for k in range(3):
        adj[:, k] = (g[:, 0] * t[0, k] + g[:, 1] * t[1, k]) + g[:, 2] * t[2, k]
    rgb1 = np.concatenate([img[:, :3], np.ones((img.shape[0], 1), np.float32)], 1)
    products = (g[:, :, None] * rgb1[:, None, :]).reshape(-1, 12).astype(np.float32)
    return adj, _tile_sums(products)


def _unorm8(b):
    return np.asarray(b, np.float32) * np.float32(1.0 / 255.0)


def score_fields(image: np.ndarray, target: np.ndarray, width: int, height: int):
    """What the loss pass leaves when scoring (train_loss.comp:34-51, 100-110):
    each pixel's mean absolute error over the channels, the view's worst in
    1/65535 (truncated), and the target's edge weight |Laplacian of the
    luminance| / 4 clamped to [0, 0.999], the luminance 0.299 r + 0.587 g +
    0.114 b added left to right, the Laplacian as differences from the centre."""
    img = np.asarray(image, np.float32).reshape(-1, 4)
    tgt = np.asarray(target, np.uint8).reshape(height, width, 4)
    y = _unorm8(tgt[..., :3]).reshape(-1, 3)
    d = np.abs(img[:, :3] - y)
    own = ((d[:, 0] + d[:, 1]).astype(np.float32) + d[:, 2]).astype(np.float32)
    error = (own / np.float32(3.0)).astype(np.float32)
    worst = int((np.clip(error, 0, 1) * np.float32(65535.0)).astype(np.float32).astype(np.uint32).max())
    lum = ((_unorm8(tgt[..., 0]) * np.float32(0.299) + _unorm8(tgt[..., 1]) * np.float32(0.587)).astype(np.float32) +
           _unorm8(tgt[..., 2]) * np.float32(0.114)).astype(np.float32)
    pad = np.pad(lum, 1, mode="edge")
    c = pad[1:-1, 1:-1]
    lap = ((((pad[1:-1, :-2] - c) + (pad[1:-1, 2:] - c)).astype(np.float32) + (pad[:-2, 1:-1] - c)).astype(np.float32) +
           (pad[2:, 1:-1] - c)).astype(np.float32)
    edge = np.clip(np.abs(lap) / np.float32(4.0), np.float32(0.0), np.float32(0.999)).astype(np.float32).reshape(-1)
    return error, worst, edge


def error_mask(error: np.ndarray, worst: int, share: float) -> np.ndarray:
    """train_error_mask.comp:14-19: 1 where the error is at least share of the
    worst (share x worst / 65535 in float32), else 0."""
    limit = (np.float32(share) * np.float32(worst)) / np.float32(65535.0)
    return np.where((worst > 0) & (np.asarray(error, np.float32) >= limit), np.float32(1.0), np.float32(0.0)).astype(np.float32)


def score_statistics(stats: np.ndarray, raw: np.ndarray, to_pixels: float, loss: float) -> np.ndarray:
    """train_score.comp:17-26: to each Gaussian's statistics (8 floats) its
    marked pixels raw[12] x to_pixels, those times the loss, and its edge
    score raw[13] x to_pixels, in float32."""
    out = np.asarray(stats, np.float32).reshape(-1, 8).copy()
    r = np.asarray(raw, np.float32).reshape(-1, 14)
    pixels = (r[:, 12] * np.float32(to_pixels)).astype(np.float32)
    out[:, 5] = out[:, 5] + pixels
    out[:, 6] = out[:, 6] + (pixels * np.float32(loss)).astype(np.float32)
    out[:, 7] = out[:, 7] + (r[:, 13] * np.float32(to_pixels)).astype(np.float32)
    return out


def gather(old: np.ndarray, source, stride: int) -> np.ndarray:
    """A per-Gaussian array of a changed cloud (train_gather.comp:16-24, used by
    trainer.cpp:749-787): element k of new Gaussian g is element k of old
    Gaussian source[g], or 0 where source[g] is negative (fresh optimiser
    state). `old` holds `stride` elements per Gaussian; returns a flat array."""
    table = np.asarray(old).reshape(-1, stride)
    src = np.asarray(source, np.int64).reshape(-1)
    if src.size and src.max() >= table.shape[0]:
        raise ValueError("a source index is past the old cloud")
    out = np.zeros((src.size, stride), table.dtype)
    keep = src >= 0
    out[keep] = table[src[keep]]
    return out.reshape(-1)


def adam_corrections(step: int):
    """1 - beta^t for Adam's step t (trainer.cpp:467-468), beta^t kept in float32
    by repeated multiplication, as the trainer keeps it on the device."""
    b1, b2 = np.float32(1.0), np.float32(1.0)
    for _ in range(step):
        b1 = np.float32(b1 * np.float32(0.9))
        b2 = np.float32(b2 * np.float32(0.999))
    return float(np.float32(1.0) - b1), float(np.float32(1.0) - b2)


def _check(code, what):
    if code != 0:
        raise RuntimeError(f"{what} failed with {code}")


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ------------------------------------------------------------------- Oracle A

@dataclass
class RefOptions:
    """ReferenceOptions (src/train/reference.hpp:18-33)."""

    background: tuple = (0.0, 0.0, 0.0)
    near_plane: float = 0.01
    no_cutoffs: bool = False
    opacity_reg: float = 0.0
    scale_reg: float = 0.0
    mip_filter_2d: bool = False
    filter3d: np.ndarray | None = None


class _Held:
    """Keeps the arrays a C struct points into alive."""

    def __init__(self):
        self.arrays = []

    def keep(self, a):
        self.arrays.append(a)
        return a


def _ccloud(cloud, held):
    arrs = [held.keep(np.ascontiguousarray(a, np.float32)) for a in
            (cloud.positions, cloud.scales, cloud.rotations, cloud.opacities, cloud.sh)]
    return CCloud(*[_p(a) for a in arrs], cloud.count, cloud.sh_degree)


def _ccamera(cam):
    c = CCamera()
    c.view[:] = [float(v) for v in cam.view_column_major()]
    c.position[:] = [float(v) for v in np.asarray(cam.position, np.float32)]
    c.fx, c.fy, c.cx, c.cy = cam.fx, cam.fy, cam.cx, cam.cy
    c.width, c.height = cam.width, cam.height
    return c


def _coptions(o, held):
    o = o or RefOptions()
    c = COptions()
    c.background[:] = list(o.background)
    c.near_plane = o.near_plane
    c.no_cutoffs = int(o.no_cutoffs)
    c.opacity_reg = o.opacity_reg
    c.scale_reg = o.scale_reg
    c.mip_filter_2d = int(o.mip_filter_2d)
    if o.filter3d is not None:
        f = held.keep(np.ascontiguousarray(o.filter3d, np.float32))
        c.filter3d = _p(f)
        c.filter3d_count = f.size
    return c


@dataclass
class RefImage:
    width: int
    height: int
    rgb: np.ndarray  # (P, 3) float64
    log_t: np.ndarray  # (P,)
    stop: np.ndarray  # (P,) uint64


def render(cloud, cam, options=None, threads=None) -> RefImage:
    held = _Held()
    P = cam.width * cam.height
    rgb = np.zeros((P, 3)); log_t = np.zeros(P); stop = np.zeros(P, np.uint64)
    cc, ca, co = _ccloud(cloud, held), _ccamera(cam), _coptions(options, held)
    _check(lib().oa_render(C.byref(cc), C.byref(ca), C.byref(co), threads or default_threads(), _p(rgb), _p(log_t), _p(stop)),
           "oa_render")
    return RefImage(cam.width, cam.height, rgb, log_t, stop)


def contributions(cloud, cam, options=None, threads=None) -> np.ndarray:
    held = _Held()
    out = np.zeros(cloud.count)
    cc, ca, co = _ccloud(cloud, held), _ccamera(cam), _coptions(options, held)
    _check(lib().oa_contributions(C.byref(cc), C.byref(ca), C.byref(co), threads or default_threads(), _p(out)),
           "oa_contributions")
    return out


def marked_pixels(cloud, cam, target, share, options=None, threads=None) -> np.ndarray:
    """reference.cpp:248-272 (FP64): per Gaussian, its composited pixels whose
    mean error is at least `share` of the view's worst."""
    held = _Held()
    out = np.zeros(cloud.count)
    tgt = np.ascontiguousarray(target, np.uint8)
    cc, ca, co = _ccloud(cloud, held), _ccamera(cam), _coptions(options, held)
    _check(lib().oa_marked_pixels(C.byref(cc), C.byref(ca), _p(tgt), float(share), C.byref(co), threads or default_threads(),
                                  _p(out)), "oa_marked_pixels")
    return out


@dataclass
class RefLoss:
    loss: float
    l1: float
    ssim: float
    adjoint: np.ndarray  # (P, 3)


def loss(image: RefImage, target: np.ndarray, lam: float, threads=None) -> RefLoss:
    rgb = np.ascontiguousarray(image.rgb, np.float64)
    tgt = np.ascontiguousarray(target, np.uint8).reshape(-1)
    adj = np.zeros_like(rgb)
    stats = np.zeros(3)
    _check(lib().oa_loss(_p(rgb), _p(tgt), image.width, image.height, lam, threads or default_threads(), _p(adj),
                         _p(stats)), "oa_loss")
    return RefLoss(float(stats[0]), float(stats[1]), float(stats[2]), adj)


@dataclass
class RefGradients:
    positions: np.ndarray
    scales: np.ndarray
    rotations: np.ndarray
    opacities: np.ndarray
    sh: np.ndarray
    absolute_centre: np.ndarray


def gradients(cloud, cam, adjoint, options=None, threads=None) -> RefGradients:
    held = _Held()
    n = cloud.count
    out = RefGradients(np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 4)), np.zeros(n),
                       np.zeros(cloud.sh.shape), np.zeros((n, 2)))
    adj = held.keep(np.ascontiguousarray(adjoint, np.float64))
    cc, ca, co = _ccloud(cloud, held), _ccamera(cam), _coptions(options, held)
    _check(lib().oa_gradients(C.byref(cc), C.byref(ca), _p(adj), C.byref(co), threads or default_threads(),
                              _p(out.positions), _p(out.scales), _p(out.rotations), _p(out.opacities), _p(out.sh),
                              _p(out.absolute_centre)), "oa_gradients")
    return out


# ------------------------------------------------------------------- Oracle B

def exp32(x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().ob_exp_array(_p(x), _p(y), x.size)
    return y


def log32(x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().ob_log_array(_p(x), _p(y), x.size)
    return y


def train_hash(index, seed: int) -> np.ndarray:
    """Oracle B's 32-bit hash behind train_uniform and the bf16 rounding."""
    idx = np.ascontiguousarray(index, np.uint32)
    out = np.empty(idx.shape, np.uint32)
    lib().ob_hash_array(_p(idx), int(seed) & 0xFFFFFFFF, idx.size, _p(out))
    return out


def train_uniform(index, seed: int) -> np.ndarray:
    """Oracle B's dropout and rounding hash (train_types.glsl:155-160): PCG's
    output permutation of index ^ (seed * 0x9E3779B9), as a float in [0, 1)."""
    idx = np.ascontiguousarray(index, np.uint32)
    out = np.empty(idx.shape, np.float32)
    lib().ob_uniform_array(_p(idx), int(seed) & 0xFFFFFFFF, idx.size, _p(out))
    return out


def bf16_stochastic(x, index, seed: int) -> np.ndarray:
    """Oracle B's bf16 stochastic rounding (bf16.glsl:16-23): the bf16 bits."""
    x = np.ascontiguousarray(x, np.float32)
    idx = np.ascontiguousarray(np.broadcast_to(np.asarray(index, np.uint32), x.shape))
    out = np.empty(x.shape, np.uint32)
    lib().ob_bf16_stochastic_array(_p(x), _p(idx), int(seed) & 0xFFFFFFFF, x.size, _p(out))
    return out


def half_stochastic(x, index, seed: int) -> np.ndarray:
    """Oracle B's fp16 stochastic rounding (train_step.glsl:70-80): the half bits."""
    x = np.ascontiguousarray(x, np.float32)
    idx = np.ascontiguousarray(np.broadcast_to(np.asarray(index, np.uint32), x.shape))
    out = np.empty(x.shape, np.uint32)
    lib().ob_half_stochastic_array(_p(x), _p(idx), int(seed) & 0xFFFFFFFF, x.size, _p(out))
    return out


def ob_camera(cam, count, sh_degree, active_degree=None, near_plane=0.01, sh_storage=0):
    """The per-view constants block, as the product's host code fills it."""
    from vsplat.fixtures import sh_words  # noqa: WPS433 (tests put the package on sys.path)
    active = sh_degree if active_degree is None else active_degree
    c = OBCamera()
    c.view[:] = [float(v) for v in cam.view_column_major()]
    c.position[:] = [float(v) for v in np.asarray(cam.position, np.float32)]
    c.near_plane = near_plane
    c.fx, c.fy, c.cx, c.cy = cam.fx, cam.fy, cam.cx, cam.cy
    c.tan_half_x = float(np.float32(0.5) * np.float32(cam.width) / np.float32(cam.fx))
    c.tan_half_y = float(np.float32(0.5) * np.float32(cam.height) / np.float32(cam.fy))
    c.width, c.height = cam.width, cam.height
    c.count = count
    c.sh_degree = active
    c.sh_words = sh_words((sh_degree + 1) ** 2, sh_storage)
    c.sh_coefficients = (active + 1) ** 2
    c.sh_format = sh_storage
    return c


@dataclass
class ModelResult:
    splats: np.ndarray  # (N, 12)
    keys: np.ndarray  # depth-sorted keys
    values: np.ndarray
    coverage: np.ndarray
    visible: int
    entries: int
    entry_keys: np.ndarray
    entry_values: np.ndarray
    ranges: np.ndarray  # (T, 2)
    image: np.ndarray | None = None  # (P, 4)
    stops: np.ndarray | None = None
    fields: np.ndarray | None = None
    adjoint: np.ndarray | None = None
    tile_totals: np.ndarray | None = None
    raw_fixed: np.ndarray | None = None
    raw: np.ndarray | None = None


class Model:
    """Oracle B over one cloud and camera: each stage as the GPU runs it."""

    def __init__(self, cloud, cam, *, near_plane=0.01, mip_filter_2d=False, filter3d=None, background=(0, 0, 0),
                 active_degree=None, sh_dropout=(0.0, 0, 0), sh_storage=0):
        from vsplat.fixtures import pack_sh
        self.cloud = cloud
        self.cam = cam
        self.n = cloud.count
        self.sh_storage = sh_storage
        self.words = pack_sh(cloud.sh, sh_storage)
        self.camera = ob_camera(cam, self.n, cloud.sh_degree, active_degree, near_plane, sh_storage)
        self.pre = OBPreOptions(int(mip_filter_2d), float(sh_dropout[0]), int(sh_dropout[1]), int(sh_dropout[2]))
        self.sh_dropout = tuple(sh_dropout)
        self.mip = mip_filter_2d
        self.filter3d = None if filter3d is None else np.ascontiguousarray(filter3d, np.float32)
        self.background = np.asarray(background, np.float32)
        self.grid_w = (cam.width + 15) // 16
        self.grid_h = (cam.height + 15) // 16
        self.tiles = self.grid_w * self.grid_h

    def fixed_scale(self):
        """3 W H 2^20: the per-pixel 1 / (3 W H) undone, then 2^20 steps per
        unit (trainer.cpp:586-591 has 64; these sums are 64-bit)."""
        return float(np.float32(3.0) * np.float32(self.cam.width) * np.float32(self.cam.height) * FIXED_STEPS)

    def forward(self) -> ModelResult:
        L = lib()
        n = self.n
        c = self.cloud
        pos = np.ascontiguousarray(c.positions, np.float32); sc = np.ascontiguousarray(c.scales, np.float32)
        rot = np.ascontiguousarray(c.rotations, np.float32); op = np.ascontiguousarray(c.opacities, np.float32)
        splats = np.zeros((n, 12), np.float32); keys = np.zeros(n, np.uint32); values = np.zeros(n, np.uint32)
        _check(L.ob_preprocess(_p(pos), _p(sc), _p(rot), _p(op), _p(self.words), _p(self.filter3d), C.byref(self.camera),
                               C.byref(self.pre), _p(splats), _p(keys), _p(values)), "ob_preprocess")
        _check(L.ob_stable_sort(_p(keys), _p(values), n), "ob_stable_sort")
        counts = np.zeros(2, np.uint64)
        coverage = np.zeros(n, np.uint32)
        _check(L.ob_bin(_p(keys), _p(values), _p(splats), n, self.grid_w, self.grid_h, _p(coverage), None, None, 0,
                        _p(counts)), "ob_bin count")
        E = int(counts[1])
        ek = np.zeros(max(E, 1), np.uint32); ev = np.zeros(max(E, 1), np.uint32)
        _check(L.ob_bin(_p(keys), _p(values), _p(splats), n, self.grid_w, self.grid_h, None, _p(ek), _p(ev), E,
                        _p(counts)), "ob_bin emit")
        ek, ev = ek[:E].copy(), ev[:E].copy()
        _check(L.ob_stable_sort(_p(ek), _p(ev), E), "ob_stable_sort entries")
        ranges = np.zeros((self.tiles, 2), np.uint32)
        _check(L.ob_tile_ranges(_p(ek), E, self.tiles, _p(ranges)), "ob_tile_ranges")
        P = self.cam.width * self.cam.height
        image = np.zeros((P, 4), np.float32); stops = np.zeros(P, np.uint32)
        _check(L.ob_forward(_p(splats), _p(ev), _p(ranges), self.cam.width, self.cam.height, _p(self.background),
                            _p(image), _p(stops)), "ob_forward")
        return ModelResult(splats, keys, values, coverage, int(counts[0]), E, ek, ev, ranges, image, stops)

    def loss(self, r: ModelResult, target, lam, *, exposure=None, score_share=None):
        """The loss and its adjoint (ob_loss). With an exposure (12 floats) the
        loss sees the exposed image and the adjoint is taken back through it,
        r.exposure_gradient holding dL/dT; with score_share the adjoint's fourth
        value marks the high-error pixels and r.edge holds the edge weights."""
        P = self.cam.width * self.cam.height
        tgt = np.ascontiguousarray(target, np.uint8).reshape(-1)
        r.fields = np.zeros((P, 9), np.float32)
        r.adjoint = np.zeros((P, 4), np.float32)
        r.tile_totals = np.zeros((self.tiles, 2), np.float32)
        seen = r.image if exposure is None else exposed_image(r.image, exposure)
        _check(lib().ob_loss(_p(np.ascontiguousarray(seen, np.float32)), _p(tgt), self.cam.width, self.cam.height, float(lam),
                             _p(r.fields), _p(r.adjoint), _p(r.tile_totals)), "ob_loss")
        if exposure is not None:
            r.adjoint, totals = exposure_backward(r.image, r.adjoint, exposure)
            # The groups' totals added in double in order (trainer.cpp:620-628).
            r.exposure_gradient = np.cumsum(totals.astype(np.float64), axis=0)[-1].astype(np.float32)
        r.edge = None
        if score_share is not None:
            error, worst, r.edge = score_fields(seen, target, self.cam.width, self.cam.height)
            r.error_max = worst
            r.adjoint = r.adjoint.copy()
            r.adjoint[:, 3] = error_mask(error, worst, score_share)
        return r

    def backward_fixed(self, r: ModelResult, adjoint=None, scale=None):
        """Fixed-point sums; scale defaults to fixed_scale(), and is
        CURVATURE_FIXED_SCALE for the curvature's +-1 adjoint."""
        adj = r.adjoint if adjoint is None else np.ascontiguousarray(adjoint, np.float32)
        scale = self.fixed_scale() if scale is None else float(scale)
        edge = getattr(r, "edge", None)
        r.raw_fixed = np.zeros((self.n, 14), np.int64)
        _check(lib().ob_backward_fixed(_p(r.splats), _p(r.entry_values), _p(r.ranges), self.cam.width, self.cam.height,
                                       _p(self.background), _p(r.image), _p(adj), _p(r.stops), scale,
                                       self.n, _p(r.raw_fixed), None if edge is None else _p(edge)), "ob_backward_fixed")
        r.raw = np.zeros((self.n, 14), np.float32)
        _check(lib().ob_fixed_to_float(_p(r.raw_fixed), self.n * 14, scale, _p(r.raw)), "ob_fixed_to_float")
        return r

    def backward_float(self, r: ModelResult, adjoint=None):
        adj = r.adjoint if adjoint is None else np.ascontiguousarray(adjoint, np.float32)
        r.raw = np.zeros((self.n, 14), np.float32)
        edge = getattr(r, "edge", None)
        _check(lib().ob_backward_float(_p(r.splats), _p(r.entry_values), _p(r.ranges), self.cam.width, self.cam.height,
                                       _p(self.background), _p(r.image), _p(adj), _p(r.stops), self.n, _p(r.raw),
                                       None if edge is None else _p(edge)), "ob_backward_float")
        return r

    def chain_options(self, mode=1, flags=1, opacity_reg=0.0, scale_reg=0.0, rates=(0, 0, 0, 0, 0), rest_scale=0.05,
                      step=1, sh_dropout=None):
        """The chain rule's options; the SH dropout defaults to the model's own
        (the product applies it in steps only: model a step with it, and
        compute_gradients without)."""
        sh_dropout = self.sh_dropout if sh_dropout is None else sh_dropout
        o = OBChainOptions()
        o.opacity_reg = float(np.float32(opacity_reg) / np.float32(self.n))
        o.scale_reg = float(np.float32(scale_reg) / np.float32(self.n))
        o.mip_filter_2d = int(self.mip)
        o.sh_dropout, o.sh_retained, o.sh_seed = float(sh_dropout[0]), int(sh_dropout[1]), int(sh_dropout[2])
        o.mode, o.flags = mode, flags
        o.rates[:] = [float(v) for v in rates]
        o.rest_scale = rest_scale
        o.beta1, o.beta2, o.epsilon = 0.9, 0.999, 1e-15
        o.correction1, o.correction2 = adam_corrections(step)
        o.seed = step
        o.sh_floats = 3 * self.cloud.coefficients
        return o

    def gradients(self, r: ModelResult, *, opacity_reg=0.0, scale_reg=0.0, stats=None):
        """Chain rule in gradients mode: (positions, scales, rotations, opacities, sh) as float32 arrays."""
        n = self.n
        c = self.cloud
        pos = np.ascontiguousarray(c.positions, np.float32).copy(); sc = np.ascontiguousarray(c.scales, np.float32).copy()
        rot = np.ascontiguousarray(c.rotations, np.float32).copy(); op = np.ascontiguousarray(c.opacities, np.float32).copy()
        words = self.words.copy()
        raw = r.raw.copy()
        st = np.zeros((n, 8), np.float32) if stats is None else stats
        states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 3 * c.coefficients)]
        g = [np.zeros(k, np.float32) for k in (n * 3, n * 3, n * 4, n, n * 3 * c.coefficients)]
        opt = self.chain_options(1, 1, opacity_reg, scale_reg)
        _check(lib().ob_chain(_p(pos), _p(sc), _p(rot), _p(op), _p(words), _p(self.filter3d), C.byref(self.camera), _p(raw),
                              _p(r.splats), _p(st), *[_p(s) for s in states], *[_p(x) for x in g], C.byref(opt)),
               "ob_chain")
        return g, st

    def pose(self, r: ModelResult) -> tuple[np.ndarray, np.ndarray]:
        """The pose gradient of a gradients-mode chain rule: the per-Gaussian
        terms (w = t x v, v = Rv dL/dp; zero where no pixel touched) and their
        total as the product sums it (``pose_totals``)."""
        n = self.n
        c = self.cloud
        pos = np.ascontiguousarray(c.positions, np.float32).copy(); sc = np.ascontiguousarray(c.scales, np.float32).copy()
        rot = np.ascontiguousarray(c.rotations, np.float32).copy(); op = np.ascontiguousarray(c.opacities, np.float32).copy()
        words = self.words.copy()
        raw = r.raw.copy()
        st = np.zeros((n, 8), np.float32)
        states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 3 * c.coefficients)]
        g = [np.zeros(k, np.float32) for k in (n * 3, n * 3, n * 4, n, n * 3 * c.coefficients)]
        terms = np.zeros((n, 6), np.float32)
        opt = self.chain_options(1, 1, 0.0, 0.0)
        opt.pose_terms = terms.ctypes.data
        _check(lib().ob_chain(_p(pos), _p(sc), _p(rot), _p(op), _p(words), _p(self.filter3d), C.byref(self.camera), _p(raw),
                              _p(r.splats), _p(st), *[_p(s) for s in states], *[_p(x) for x in g], C.byref(opt)),
               "ob_chain")
        return terms, pose_totals(terms)

    def statistics(self, r: ModelResult, stats=None) -> np.ndarray:
        """Chain rule in statistics mode (collect_statistics, score_view): only
        the densification statistics, added to `stats` (8 floats per Gaussian)."""
        n = self.n
        c = self.cloud
        pos = np.ascontiguousarray(c.positions, np.float32).copy(); sc = np.ascontiguousarray(c.scales, np.float32).copy()
        rot = np.ascontiguousarray(c.rotations, np.float32).copy(); op = np.ascontiguousarray(c.opacities, np.float32).copy()
        words = self.words.copy()
        raw = r.raw.copy()
        st = np.zeros((n, 8), np.float32) if stats is None else stats
        states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 3 * c.coefficients)]
        opt = self.chain_options(0, 1)
        _check(lib().ob_chain(_p(pos), _p(sc), _p(rot), _p(op), _p(words), _p(self.filter3d), C.byref(self.camera), _p(raw),
                              _p(r.splats), _p(st), *[_p(x) for x in states], None, None, None, None, None, C.byref(opt)),
               "ob_chain")
        return st

    def step(self, r: ModelResult, rates, *, rest_scale=0.05, step=1, selective=True, opacity_reg=0.0, scale_reg=0.0, states=None,
             second_order=False, trust_region=0.0, curvature_updates=1):
        """Chain rule in step mode, then SH's part of the step (train_gradient.comp,
        train_sh_step.comp), from zero optimiser state or from `states` (five
        uint32 arrays, updated in place): the parameters after Adam's step
        number `step`, or with second_order the trust-region step of
        train_step.glsl:44-61, as (positions, scales, rotations, opacities, sh_words)."""
        n = self.n
        c = self.cloud
        pos = np.ascontiguousarray(c.positions, np.float32).reshape(-1).copy()
        sc = np.ascontiguousarray(c.scales, np.float32).reshape(-1).copy()
        rot = np.ascontiguousarray(c.rotations, np.float32).reshape(-1).copy()
        op = np.ascontiguousarray(c.opacities, np.float32).copy()
        words = self.words.copy()
        raw = r.raw.copy()
        st = np.zeros((n, 8), np.float32)
        if states is None:
            states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 3 * c.coefficients)]
        flags = (4 if selective else 0) | (16 if second_order else 0)
        opt = self.chain_options(2, flags, opacity_reg, scale_reg, rates=rates, rest_scale=rest_scale, step=step)
        opt.trust_region = float(trust_region)
        opt.curvature_scale = curvature_scale(curvature_updates)
        opt.curvature_beta = float(CURVATURE_BETA)
        _check(lib().ob_chain(_p(pos), _p(sc), _p(rot), _p(op), _p(words), _p(self.filter3d), C.byref(self.camera), _p(raw),
                              _p(r.splats), _p(st), *[_p(s) for s in states], None, None, None, None, None, C.byref(opt)),
               "ob_chain")
        _check(lib().ob_sh_step(_p(words), _p(states[4]), _p(raw), C.byref(self.camera), C.byref(opt)), "ob_sh_step")
        return pos, sc, rot, op, words

    def curvature(self, r: ModelResult, states, updates: int):
        """Trainer::update_curvature's chain (train_gradient.comp and
        train_sh_step.comp in curvature mode): from the raw sums of the
        backward pass of a Hutchinson adjoint (hutchinson_adjoint), each
        parameter's squared gradient per image value into the running average
        in its state's second half; rotations keep Adam. `updates` is the
        count including this one (the rounding seed)."""
        c = self.cloud
        pos = np.ascontiguousarray(c.positions, np.float32).reshape(-1).copy()
        sc = np.ascontiguousarray(c.scales, np.float32).reshape(-1).copy()
        rot = np.ascontiguousarray(c.rotations, np.float32).reshape(-1).copy()
        op = np.ascontiguousarray(c.opacities, np.float32).copy()
        words = self.words.copy()
        raw = r.raw.copy()
        st = np.zeros((self.n, 8), np.float32)
        opt = self.chain_options(3, 0)
        opt.seed = int(updates)
        opt.curvature_beta = float(CURVATURE_BETA)
        opt.per_value = float(np.float32(1.0) / np.float32(3.0 * self.cam.width * self.cam.height))
        _check(lib().ob_chain(_p(pos), _p(sc), _p(rot), _p(op), _p(words), _p(self.filter3d), C.byref(self.camera), _p(raw),
                              _p(r.splats), _p(st), *[_p(s) for s in states], None, None, None, None, None, C.byref(opt)),
               "ob_chain")
        _check(lib().ob_sh_step(_p(words), _p(states[4]), _p(raw), C.byref(self.camera), C.byref(opt)), "ob_sh_step")
        return states


def pose_totals(terms: np.ndarray) -> np.ndarray:
    """The pose gradient from per-Gaussian terms (trainer.cpp:630-637 with
    train_gradient.comp:323-337): per group of 256 Gaussians a float32 tree
    (element l += element l + s for s = 128, 64, ..., 1), then the groups'
    totals summed in double in order and rounded to float."""
    t = np.ascontiguousarray(terms, np.float32).reshape(-1, 6)
    groups = (t.shape[0] + 255) // 256
    x = np.zeros((groups * 256, 6), np.float32)
    x[: t.shape[0]] = t
    x = x.reshape(groups, 256, 6)
    s = 128
    while s >= 1:
        x[:, :s] = x[:, :s] + x[:, s:2 * s]
        s //= 2
    total = np.zeros(6, np.float64)
    for row in x[:, 0].astype(np.float64):
        total = total + row
    return total.astype(np.float32)


def refined_pose(cam, move):
    """trainer.cpp:120-137: the camera moved by (w, v), a rotation w (axis times
    angle, Rodrigues' formula with its small-angle limits) applied after the
    view and then a translation v, in camera space, in float64; the view and
    position rounded to float32."""
    w = np.asarray([float(np.float32(m)) for m in move[:3]], np.float64)
    v = np.asarray([float(np.float32(m)) for m in move[3:6]], np.float64)
    angle = float(np.linalg.norm(w))
    K = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    a = math.sin(angle) / angle if angle > 1e-8 else 1.0
    b = (1.0 - math.cos(angle)) / (angle * angle) if angle > 1e-8 else 0.5
    R = np.eye(3) + a * K + b * (K @ K)
    before = np.asarray(cam.view, np.float32).astype(np.float64)
    rotation = R @ before[:3, :3]
    translation = R @ before[:3, 3] + v
    after = np.eye(4)
    after[:3, :3] = rotation
    after[:3, 3] = translation
    out = type(cam)(**{k: getattr(cam, k) for k in ("fx", "fy", "cx", "cy", "width", "height")})
    out.view = after.astype(np.float32)
    out.position = (-(rotation.T @ translation)).astype(np.float32)
    return out


def _hash(x):
    """train_hutchinson.comp:14-21, in uint64 reduced to 32 bits after each product."""
    x = np.atleast_1d(np.asarray(x, np.uint64)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    return x.astype(np.uint32)


def hutchinson_adjoint(pixels: int, seed: int) -> np.ndarray:
    """train_hutchinson.comp:15-29: per pixel and channel a random sign from
    hash(i ^ hash(seed)), bits 0, 1, 2; the fourth value 0."""
    bits = _hash(np.arange(pixels, dtype=np.uint32) ^ _hash(seed)[0])
    out = np.zeros((pixels, 4), np.float32)
    for ch in range(3):
        out[:, ch] = np.where((bits >> np.uint32(ch)) & np.uint32(1), np.float32(1.0), np.float32(-1.0))
    return out


_M32 = np.uint64(0xFFFFFFFF)


def pcg_hash(v):
    """train_noise.comp:23-28: PCG's output permutation of a 32-bit state, in
    uint64 reduced to 32 bits after each product."""
    v = np.atleast_1d(np.asarray(v, np.uint64)) & _M32
    state = (v * np.uint64(747796405) + np.uint64(2891336453)) & _M32
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & _M32
    return ((word >> np.uint64(22)) ^ word).astype(np.uint32)


def noise_uniforms(count: int, iteration: int) -> np.ndarray:
    """train_noise.comp:31-33: four uniforms in (0, 1] per Gaussian, float32 as
    the shader rounds them: float(hash) + 1, then / 4294967297 (which float32
    holds as 2^32)."""
    i = np.arange(count, dtype=np.uint64)
    salt = np.uint64((int(iteration) * 0x9E3779B9) & 0xFFFFFFFF)
    out = np.empty((count, 4), np.float32)
    for k in range(4):
        h = pcg_hash(pcg_hash((i * np.uint64(4) + np.uint64(k)) & _M32).astype(np.uint64) ^ salt)
        out[:, k] = (h.astype(np.float32) + np.float32(1.0)) / np.float32(4294967297.0)
    return out


def position_noise(positions, scales, rotations, opacities, scaler: float, iteration: int) -> np.ndarray:
    """train_noise.comp:35-70 in float64 from the shader's float32 uniforms:
    positions + scaler * gate * Sigma n, gate = logistic(100 ((1 - opacity) -
    0.995)), n three Box-Muller normals, Sigma = R diag(exp(2 s)) R^T of the
    normalised rotation; Gaussians whose gate is below 1e-6 or whose move is
    not finite stay where they are."""
    p = np.asarray(positions, np.float64).reshape(-1, 3)
    n_ = p.shape[0]
    o = 1.0 / (1.0 + np.exp(-np.asarray(opacities, np.float64).reshape(-1)))
    with np.errstate(over="ignore"):
        gate = 1.0 / (1.0 + np.exp(-100.0 * ((1.0 - o) - 0.995)))
    u = noise_uniforms(n_, iteration).astype(np.float64)
    r0, r1 = np.sqrt(-2.0 * np.log(u[:, 0])), np.sqrt(-2.0 * np.log(u[:, 2]))
    two_pi = 2.0 * np.pi
    normal = np.stack([r0 * np.cos(two_pi * u[:, 1]), r0 * np.sin(two_pi * u[:, 1]), r1 * np.cos(two_pi * u[:, 3])], 1)
    q = np.asarray(rotations, np.float64).reshape(-1, 4).copy()
    length = np.linalg.norm(q, axis=1)
    q[length > 0] /= length[length > 0, None]
    q[length <= 0] = [1.0, 0.0, 0.0, 0.0]
    w, x, y, z = q.T
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
                  np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
                  np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)
    s2 = np.exp(2.0 * np.asarray(scales, np.float64).reshape(-1, 3))
    sigma = np.einsum("nij,nj,nkj->nik", R, s2, R)
    with np.errstate(over="ignore", invalid="ignore"):
        move = float(scaler) * gate[:, None] * np.einsum("nij,nj->ni", sigma, normal)
    keep = (gate >= 1e-6) & np.all(np.isfinite(move), axis=1)
    out = p.copy()
    out[keep] += move[keep]
    return out


def opacity_limit(maximum: float) -> float:
    """trainer.cpp:811: the logit log(m / (1 - m)) that reset_opacity caps
    stored opacities at, in float64."""
    m = float(np.float32(maximum))
    return float(np.log(m / (1.0 - m)))


def adam_from_gradients(cloud, words, grads, rates, *, rest_scale=0.05, step=1, selective=True, sh_storage=0, states=None):
    """The batched step's Adam (ob_adam): the cloud's parameters (SH as packed
    words) after Adam's step number `step` with the summed gradients grads =
    (positions, scales, rotations, opacities, sh) as flat float32 arrays, from
    zero optimiser state or from `states` (five uint32 arrays, one word per
    parameter, updated in place)."""
    n = cloud.count
    pos = np.ascontiguousarray(cloud.positions, np.float32).reshape(-1).copy()
    sc = np.ascontiguousarray(cloud.scales, np.float32).reshape(-1).copy()
    rot = np.ascontiguousarray(cloud.rotations, np.float32).reshape(-1).copy()
    op = np.ascontiguousarray(cloud.opacities, np.float32).copy()
    w = np.ascontiguousarray(words, np.uint32).copy()
    if states is None:
        states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 3 * cloud.coefficients)]
    for s, k in zip(states, (n * 3, n * 3, n * 4, n, n * 3 * cloud.coefficients)):
        if s.dtype != np.uint32 or s.size != k or not s.flags.c_contiguous:
            raise ValueError("states: five contiguous uint32 arrays, one word per parameter")
    g = [np.ascontiguousarray(x, np.float32).reshape(-1) for x in grads]
    opt = OBChainOptions()
    opt.mode, opt.flags = 2, 4 if selective else 0
    opt.rates[:] = [float(v) for v in rates]
    opt.rest_scale = rest_scale
    opt.beta1, opt.beta2, opt.epsilon = 0.9, 0.999, 1e-15
    opt.correction1, opt.correction2 = adam_corrections(step)
    opt.seed = step
    opt.sh_floats = 3 * cloud.coefficients
    _check(lib().ob_adam(_p(pos), _p(sc), _p(rot), _p(op), _p(w), *[_p(s) for s in states], *[_p(x) for x in g], n, w.shape[1],
                         sh_storage, C.byref(opt)), "ob_adam")
    return pos, sc, rot, op, w
