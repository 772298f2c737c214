"""vsplat: Python access to libvsplat.so, the MI355X differentiable Gaussian-splat
tile rasterizer, through its C interface (include/vsplat.h).

``Trainer`` mirrors the reference's ``class Trainer`` (src/train/trainer.hpp:89-270):
``render``, ``compute_gradients``, ``step``, ``apply_gradients``, ``gradients``,
``cloud`` and the rest, with the reference's errors as Python exceptions
(``RuntimeError``, ``BudgetExceeded``, ``ValueError`` for std::invalid_argument,
``LogicError`` for std::logic_error). There is no fallback: without the built
library or a GPU, constructing a Trainer raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import fixtures  # noqa: F401  (re-exported)
from .fixtures import Camera, Cloud

LIBRARY = Path(__file__).resolve().parent / "libvsplat.so"

OK, RUNTIME_ERROR, BUDGET_EXCEEDED, INVALID_ARGUMENT, LOGIC_ERROR, DEVICE_ERROR, UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6
BYTES_PER_TILE_ENTRY = 16
MODE_STATISTICS, MODE_GRADIENTS, MODE_STEP = 0, 1, 2
SH_HALF, SH_BF16 = 0, 1  # TrainerOptions.sh_storage
TILES_16X16, TILES_8X8, TILES_32X8, TILES_AUTO, TILES_32X32 = 0, 1, 2, 3, 4  # TrainerOptions.tile_shape
BATCH_VIEW_STREAMS, BATCH_ONE_LAUNCH = 0, 1  # TrainerOptions.batch_raster
# vsplat_render_noalloc's limits (VSPLAT_NOALLOC_MAX_*, include/vsplat.h).
NOALLOC_MAX_GAUSSIANS, NOALLOC_MAX_ENTRIES, NOALLOC_MAX_PIXELS, NOALLOC_MAX_TILES = 1 << 17, 1 << 21, 1 << 21, 1 << 16

EXPORTS = [
    "vsplat_default_options", "vsplat_default_rates", "vsplat_trainer_create", "vsplat_trainer_destroy", "vsplat_render",
    "vsplat_render_noalloc",
    "vsplat_compute_gradients", "vsplat_collect_statistics", "vsplat_apply_gradients", "vsplat_step",
    "vsplat_prefetch_target", "vsplat_get_gradients", "vsplat_get_cloud", "vsplat_get_densify_stats",
    "vsplat_reset_densify_stats", "vsplat_set_active_sh_degree", "vsplat_set_filter3d", "vsplat_set_sh_dropout",
    "vsplat_parameters", "vsplat_gaussian_count", "vsplat_sh_degree", "vsplat_active_sh_degree",
    "vsplat_uses_fixed_point", "vsplat_gpu_bytes", "vsplat_bytes_per_gaussian", "vsplat_staging_bytes",
    "vsplat_timings", "vsplat_last_error", "vsplat_enqueue", "vsplat_synchronize", "vsplat_stream",
    "vsplat_reserve_entries", "vsplat_inspect", "vsplat_inspect_view", "vsplat_math_probe", "vsplat_reset_timings",
    "vsplat_set_stage_timing", "vsplat_step_views", "vsplat_enqueue_views", "vsplat_synchronize_views",
    "vsplat_capture_views", "vsplat_launch_captured", "vsplat_release_captured",
    "vsplat_apply_changes", "vsplat_device_alloc", "vsplat_device_read", "vsplat_device_copy", "vsplat_device_free",
    "vsplat_autotune_tiles", "vsplat_autotune_views", "vsplat_score_view", "vsplat_set_exposure",
    "vsplat_ply_info", "vsplat_read_ply", "vsplat_write_ply", "vsplat_save_ply",
    "vsplat_enable_second_order", "vsplat_update_curvature", "vsplat_reset_opacity", "vsplat_add_position_noise",
    "vsplat_refined_pose",
]


class BudgetExceeded(RuntimeError):
    """splat::BudgetExceeded (src/train/trainer.hpp:46-49): recoverable."""


class LogicError(Exception):
    """std::logic_error."""


class Unsupported(RuntimeError):
    """An option this build does not implement."""


class DeviceError(RuntimeError):
    """A HIP call failed, or there is no GPU."""


class CCloud(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("scales", C.c_void_p), ("rotations", C.c_void_p), ("opacities", C.c_void_p),
                ("sh", C.c_void_p), ("count", C.c_uint32), ("sh_degree", C.c_int32)]


class CCamera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("position", C.c_float * 3), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("width", C.c_int32), ("height", C.c_int32)]


class COptions(C.Structure):
    _fields_ = [("lambda_ssim", C.c_float), ("background", C.c_float * 3), ("near_plane", C.c_float),
                ("force_fixed_point", C.c_int32), ("selective_adam", C.c_int32), ("opacity_reg", C.c_float),
                ("scale_reg", C.c_float), ("position_noise", C.c_float), ("mip_filter_2d", C.c_int32),
                ("pose_gradients", C.c_int32), ("memory_budget", C.c_uint64), ("sh_storage", C.c_int32),
                ("recompute_transmittance", C.c_int32), ("tile_shape", C.c_int32), ("batch_raster", C.c_int32)]


class CRates(C.Structure):
    _fields_ = [("position", C.c_float), ("scale", C.c_float), ("rotation", C.c_float), ("opacity", C.c_float),
                ("sh", C.c_float), ("sh_rest_scale", C.c_float), ("trust_region", C.c_float)]


class CStats(C.Structure):
    _fields_ = [("loss", C.c_float), ("l1", C.c_float), ("ssim", C.c_float), ("visible_splats", C.c_uint32),
                ("tile_entries", C.c_uint32), ("exposure_gradient", C.c_float * 12), ("pose_gradient", C.c_float * 6)]


class CTiming(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("milliseconds", C.c_double)]


_LIB = None


def lib():
    """The loaded library; raises ImportError when it is not built."""
    global _LIB
    if _LIB is None:
        if not LIBRARY.exists():
            raise ImportError(f"{LIBRARY} is not built (run __graft_entry__.build() or make -C vibe-splat_amd)")
        L = C.CDLL(str(LIBRARY))
        vp = C.c_void_p
        L.vsplat_default_options.argtypes = [vp]
        L.vsplat_default_options.restype = None
        L.vsplat_default_rates.argtypes = [vp]
        L.vsplat_default_rates.restype = None
        L.vsplat_trainer_create.argtypes = [vp, vp, vp]
        L.vsplat_trainer_destroy.argtypes = [vp]
        L.vsplat_trainer_destroy.restype = None
        L.vsplat_render.argtypes = [vp, vp, vp]
        L.vsplat_render_noalloc.argtypes = [vp, vp, vp, C.c_uint64]
        L.vsplat_compute_gradients.argtypes = [vp, vp, vp, C.c_size_t, vp]
        L.vsplat_collect_statistics.argtypes = [vp, vp, vp, C.c_size_t, vp]
        L.vsplat_apply_gradients.argtypes = [vp, vp]
        L.vsplat_step.argtypes = [vp, vp, vp, C.c_size_t, vp, vp, C.c_size_t, vp]
        L.vsplat_prefetch_target.argtypes = [vp, vp, C.c_size_t]
        L.vsplat_get_gradients.argtypes = [vp] * 6
        L.vsplat_get_cloud.argtypes = [vp] * 6
        L.vsplat_get_densify_stats.argtypes = [vp, vp]
        L.vsplat_reset_densify_stats.argtypes = [vp]
        L.vsplat_set_active_sh_degree.argtypes = [vp, C.c_int32]
        L.vsplat_set_filter3d.argtypes = [vp, vp, C.c_size_t]
        L.vsplat_set_sh_dropout.argtypes = [vp, C.c_float, C.c_int32, C.c_uint32]
        L.vsplat_parameters.argtypes = [vp, C.c_uint32, vp, vp]
        for name, rt in (("vsplat_gaussian_count", C.c_uint32), ("vsplat_sh_degree", C.c_int32),
                         ("vsplat_active_sh_degree", C.c_int32), ("vsplat_uses_fixed_point", C.c_int32),
                         ("vsplat_gpu_bytes", C.c_uint64), ("vsplat_bytes_per_gaussian", C.c_uint64),
                         ("vsplat_staging_bytes", C.c_uint64)):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = rt
        L.vsplat_timings.argtypes = [vp, vp, C.c_uint32, vp]
        L.vsplat_last_error.argtypes = []
        L.vsplat_last_error.restype = C.c_char_p
        L.vsplat_enqueue.argtypes = [vp, vp, vp, C.c_int32, vp]
        L.vsplat_synchronize.argtypes = [vp, vp]
        L.vsplat_stream.argtypes = [vp]
        L.vsplat_stream.restype = vp
        L.vsplat_reserve_entries.argtypes = [vp, C.c_uint64]
        L.vsplat_inspect.argtypes = [vp, C.c_char_p, vp, C.c_uint64, vp]
        L.vsplat_inspect_view.argtypes = [vp, C.c_uint32, C.c_char_p, vp, C.c_uint64, vp]
        L.vsplat_math_probe.argtypes = [C.c_int32, vp, vp, C.c_uint64]
        L.vsplat_reset_timings.argtypes = [vp]
        L.vsplat_set_stage_timing.argtypes = [vp, C.c_int32]
        L.vsplat_step_views.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_int32, vp, vp]
        L.vsplat_enqueue_views.argtypes = [vp, C.c_uint32, vp, vp, C.c_int32, vp]
        L.vsplat_synchronize_views.argtypes = [vp, vp, C.c_uint32, vp]
        L.vsplat_capture_views.argtypes = [vp, C.c_uint32, vp, vp, C.c_int32, vp]
        L.vsplat_launch_captured.argtypes = [vp, C.c_uint32]
        L.vsplat_release_captured.argtypes = [vp]
        L.vsplat_device_alloc.argtypes = [C.c_uint64, vp, vp]
        L.vsplat_device_read.argtypes = [vp, vp, C.c_uint64]
        L.vsplat_apply_changes.argtypes = [vp, vp, vp, C.c_uint32]
        L.vsplat_device_copy.argtypes = [vp, vp, C.c_uint64, C.c_uint32, vp]
        L.vsplat_device_free.argtypes = [vp]
        L.vsplat_device_free.restype = None
        L.vsplat_autotune_tiles.argtypes = [vp, vp, vp, C.c_uint32, vp, vp]
        L.vsplat_autotune_views.argtypes = [vp, C.c_uint32, vp, vp, C.c_uint32, vp, vp]
        L.vsplat_score_view.argtypes = [vp, vp, vp, C.c_size_t, C.c_float, C.c_float, vp]
        L.vsplat_set_exposure.argtypes = [vp, vp]
        L.vsplat_enable_second_order.argtypes = [vp]
        L.vsplat_update_curvature.argtypes = [vp]
        L.vsplat_reset_opacity.argtypes = [vp, C.c_float]
        L.vsplat_add_position_noise.argtypes = [vp, C.c_float, C.c_uint32]
        L.vsplat_ply_info.argtypes = [C.c_char_p, vp, vp]
        L.vsplat_refined_pose.argtypes = [vp, vp, vp]
        L.vsplat_read_ply.argtypes = [C.c_char_p, vp, vp, vp, vp, vp, C.c_uint64]
        L.vsplat_write_ply.argtypes = [C.c_char_p, vp]
        L.vsplat_save_ply.argtypes = [vp, C.c_char_p]
        _LIB = L
    return _LIB


def _raise(code: int, what: str):
    if code == OK:
        return
    message = f"{what}: {lib().vsplat_last_error().decode(errors='replace')}"
    if code == BUDGET_EXCEEDED:
        raise BudgetExceeded(message)
    if code == INVALID_ARGUMENT:
        raise ValueError(message)
    if code == LOGIC_ERROR:
        raise LogicError(message)
    if code == UNSUPPORTED:
        raise Unsupported(message)
    if code == DEVICE_ERROR:
        raise DeviceError(message)
    raise RuntimeError(message)


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


@dataclass
class TrainerOptions:
    """TrainerOptions (src/train/trainer.hpp:24-42)."""

    lambda_ssim: float = 0.2
    background: tuple = (0.0, 0.0, 0.0)
    near_plane: float = 0.01
    force_fixed_point: bool = False
    selective_adam: bool = True
    opacity_reg: float = 0.0
    scale_reg: float = 0.0
    position_noise: float = 0.0
    mip_filter_2d: bool = False
    pose_gradients: bool = False
    memory_budget: int = 2 ** 64 - 1
    sh_storage: int = 0  # SH_HALF (the reference's layout) or SH_BF16 (this build only)
    recompute_transmittance: bool = False  # backward without the forward's stored log T and stops (this build only)
    tile_shape: int = 0  # TILES_16X16 (the reference's), TILES_8X8, TILES_32X8, TILES_32X32 or TILES_AUTO (this build only)
    batch_raster: int = 0  # BATCH_VIEW_STREAMS or BATCH_ONE_LAUNCH: how a batch of views runs (this build only)

    def to_c(self) -> COptions:
        o = COptions()
        o.lambda_ssim = self.lambda_ssim
        o.background[:] = [float(v) for v in self.background]
        o.near_plane = self.near_plane
        o.force_fixed_point = int(self.force_fixed_point)
        o.selective_adam = int(self.selective_adam)
        o.opacity_reg, o.scale_reg, o.position_noise = self.opacity_reg, self.scale_reg, self.position_noise
        o.mip_filter_2d = int(self.mip_filter_2d)
        o.pose_gradients = int(self.pose_gradients)
        o.memory_budget = int(self.memory_budget)
        o.sh_storage = int(self.sh_storage)
        o.recompute_transmittance = int(self.recompute_transmittance)
        o.tile_shape = int(self.tile_shape)
        o.batch_raster = int(self.batch_raster)
        return o


@dataclass
class LearningRates:
    """LearningRates (src/train/trainer.hpp:51-61)."""

    position: float = 1.6e-4
    scale: float = 0.005
    rotation: float = 0.001
    opacity: float = 0.05
    sh: float = 0.025
    sh_rest_scale: float = 0.05
    trust_region: float = 0.0

    def to_c(self) -> CRates:
        return CRates(self.position, self.scale, self.rotation, self.opacity, self.sh, self.sh_rest_scale, self.trust_region)


@dataclass
class StepStats:
    """StepStats (src/train/trainer.hpp:63-73)."""

    loss: float
    l1: float
    ssim: float
    visible_splats: int
    tile_entries: int
    exposure_gradient: tuple = (0.0,) * 12
    pose_gradient: tuple = (0.0,) * 6

    @staticmethod
    def of(c: CStats) -> "StepStats":
        return StepStats(c.loss, c.l1, c.ssim, c.visible_splats, c.tile_entries, tuple(c.exposure_gradient), tuple(c.pose_gradient))


@dataclass
class Gradients:
    positions: np.ndarray
    scales: np.ndarray
    rotations: np.ndarray
    opacities: np.ndarray
    sh: np.ndarray


def camera_to_c(cam: Camera) -> CCamera:
    c = CCamera()
    c.view[:] = [float(v) for v in cam.view_column_major()]
    c.position[:] = [float(v) for v in np.asarray(cam.position, np.float32)]
    c.fx, c.fy, c.cx, c.cy = cam.fx, cam.fy, cam.cx, cam.cy
    c.width, c.height = cam.width, cam.height
    return c


INSPECT_TYPES = {"splats": (np.float32, 12), "keys": (np.uint32, 1), "values": (np.uint32, 1), "coverage": (np.uint32, 1),
                 "entry_keys": (np.uint32, 1), "entry_blocks": (np.uint32, 1), "entry_values": (np.uint32, 1), "ranges": (np.uint32, 2),
                 "image": (np.float32, 4), "stops": (np.uint32, 1), "adjoint": (np.float32, 4), "fields": (np.float32, 9),
                 "tile_totals": (np.float32, 2), "block_work": (np.uint32, 1), "block_order": (np.uint32, 1), "list_order": (np.uint32, 1), "counters": (np.uint32, 1),
                 "state_positions": (np.uint32, 3), "state_scales": (np.uint32, 3), "state_rotations": (np.uint32, 4),
                 "state_opacities": (np.uint32, 1), "state_sh": (np.uint32, 1), "stats": (np.float32, 8)}


class Trainer:
    """The reference's Trainer over the C interface (src/train/trainer.hpp:89-270)."""

    def __init__(self, cloud: Cloud, options: TrainerOptions | None = None):
        self._lib = lib()
        self.options = options or TrainerOptions()
        self._shape = cloud
        arrays = [np.ascontiguousarray(a, np.float32) for a in
                  (cloud.positions, cloud.scales, cloud.rotations, cloud.opacities, cloud.sh)]
        cc = CCloud(*[_p(a) for a in arrays], cloud.count, cloud.sh_degree)
        handle = C.c_void_p()
        opts = self.options.to_c()
        _raise(self._lib.vsplat_trainer_create(C.byref(cc), C.byref(opts), C.byref(handle)), "Trainer")
        self._h = handle
        self.count = cloud.count
        self.sh_degree = cloud.sh_degree
        self._next = None  # keeps a prefetched target alive until the next step

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vsplat_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 (interpreter shutdown)
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- the per-step surface
    def render(self, camera: Camera) -> np.ndarray:
        out = np.zeros((camera.height * camera.width, 4), np.float32)
        cam = camera_to_c(camera)
        _raise(self._lib.vsplat_render(self._h, C.byref(cam), _p(out)), "render")
        return out

    def render_noalloc(self, camera: Camera, out: "DeviceBuffer") -> None:
        """vsplat_render_noalloc: the render into GPU memory (out), with no
        allocation; the image is out's first width*height*16 bytes."""
        cam = camera_to_c(camera)
        _raise(self._lib.vsplat_render_noalloc(self._h, C.byref(cam), C.c_void_p(out.ptr), out.nbytes), "render_noalloc")

    def compute_gradients(self, camera: Camera, target: np.ndarray) -> StepStats:
        t = np.ascontiguousarray(target, np.uint8)
        cam, st = camera_to_c(camera), CStats()
        _raise(self._lib.vsplat_compute_gradients(self._h, C.byref(cam), _p(t), t.nbytes, C.byref(st)), "compute_gradients")
        return StepStats.of(st)

    def collect_statistics(self, camera: Camera, target: np.ndarray) -> StepStats:
        t = np.ascontiguousarray(target, np.uint8)
        cam, st = camera_to_c(camera), CStats()
        _raise(self._lib.vsplat_collect_statistics(self._h, C.byref(cam), _p(t), t.nbytes, C.byref(st)), "collect_statistics")
        return StepStats.of(st)

    def score_view(self, camera: Camera, target: np.ndarray, error_share: float, pixel_weight: float = 1.0) -> StepStats:
        """Trainer::score_view (trainer.cpp:519-532): marked pixels, marked loss
        and edge score added to the densification statistics."""
        t = np.ascontiguousarray(target, np.uint8)
        cam, st = camera_to_c(camera), CStats()
        _raise(self._lib.vsplat_score_view(self._h, C.byref(cam), _p(t), t.nbytes, error_share, pixel_weight, C.byref(st)),
               "score_view")
        return StepStats.of(st)

    def save_ply(self, path):
        """The trainer's cloud written as a 3DGS .ply."""
        _raise(self._lib.vsplat_save_ply(self._h, str(path).encode()), "save_ply")

    def enable_second_order(self):
        """Trainer::enable_second_order (trainer.hpp:117)."""
        _raise(self._lib.vsplat_enable_second_order(self._h), "enable_second_order")

    def update_curvature(self):
        """Trainer::update_curvature (trainer.hpp:118)."""
        _raise(self._lib.vsplat_update_curvature(self._h), "update_curvature")

    def reset_opacity(self, maximum):
        """Trainer::reset_opacity (trainer.hpp:163): opacities capped at `maximum`, their optimiser state cleared."""
        _raise(self._lib.vsplat_reset_opacity(self._h, float(maximum)), "reset_opacity")

    def add_position_noise(self, scaler, iteration):
        """Trainer::add_position_noise (trainer.hpp:165): one round of 3DGS-MCMC's position noise."""
        _raise(self._lib.vsplat_add_position_noise(self._h, float(scaler), int(iteration)), "add_position_noise")

    def set_exposure(self, transform):
        """Trainer::set_exposure (trainer.hpp:154): 12 floats, rows of (r, g, b, 1)."""
        t = np.ascontiguousarray(transform, np.float32).reshape(-1)
        if t.size != 12:
            raise ValueError("an exposure is 12 floats")
        _raise(self._lib.vsplat_set_exposure(self._h, _p(t)), "set_exposure")

    def step(self, camera: Camera, target: np.ndarray, rates: LearningRates, next_target: np.ndarray | None = None) -> StepStats:
        t = np.ascontiguousarray(target, np.uint8)
        n = None if next_target is None else np.ascontiguousarray(next_target, np.uint8)
        cam, st, r = camera_to_c(camera), CStats(), rates.to_c()
        code = self._lib.vsplat_step(self._h, C.byref(cam), _p(t), t.nbytes, C.byref(r), _p(n), 0 if n is None else n.nbytes,
                                     C.byref(st))
        self._next = n
        _raise(code, "step")
        return StepStats.of(st)

    def prefetch_target(self, target: np.ndarray):
        t = np.ascontiguousarray(target, np.uint8)
        self._next = t
        _raise(self._lib.vsplat_prefetch_target(self._h, _p(t), t.nbytes), "prefetch_target")

    def apply_gradients(self, rates: LearningRates):
        r = rates.to_c()
        _raise(self._lib.vsplat_apply_gradients(self._h, C.byref(r)), "apply_gradients")

    # --- results
    def gradients(self) -> Gradients:
        n = self.count
        coefficients = (self.sh_degree + 1) ** 2
        g = Gradients(np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32),
                      np.zeros(n, np.float32), np.zeros((n, coefficients, 3), np.float32))
        _raise(self._lib.vsplat_get_gradients(self._h, _p(g.positions), _p(g.scales), _p(g.rotations), _p(g.opacities), _p(g.sh)),
               "gradients")
        return g

    def cloud(self) -> Cloud:
        c = Cloud.zeros(self.count, self.sh_degree)
        _raise(self._lib.vsplat_get_cloud(self._h, _p(c.positions), _p(c.scales), _p(c.rotations), _p(c.opacities), _p(c.sh)),
               "cloud")
        return c

    def densify_stats(self) -> np.ndarray:
        out = np.zeros((self.count, 8), np.float32)
        _raise(self._lib.vsplat_get_densify_stats(self._h, _p(out)), "densify_stats")
        return out

    def reset_densify_stats(self):
        _raise(self._lib.vsplat_reset_densify_stats(self._h), "reset_densify_stats")

    def apply_changes(self, cloud: Cloud | None, source) -> None:
        """Trainer::apply_changes (trainer.cpp:749-787): the Gaussians become
        len(source) new ones; new Gaussian g keeps old Gaussian source[g]'s
        optimiser state (fresh when negative) and takes its values from
        `cloud`, or from old Gaussian source[g] when cloud is None."""
        src = np.ascontiguousarray(source, np.int64)
        count = src.size
        cc = None
        if cloud is not None:
            arrays = [np.ascontiguousarray(a, np.float32) for a in
                      (cloud.positions, cloud.scales, cloud.rotations, cloud.opacities, cloud.sh)]
            cc = CCloud(*[_p(a) for a in arrays], cloud.count, cloud.sh_degree)
        _raise(self._lib.vsplat_apply_changes(self._h, None if cc is None else C.byref(cc), _p(src), count), "apply_changes")
        self.count = count

    # --- settings
    def set_active_sh_degree(self, degree: int):
        _raise(self._lib.vsplat_set_active_sh_degree(self._h, degree), "set_active_sh_degree")

    def set_filter3d(self, values: np.ndarray):
        f = np.ascontiguousarray(values, np.float32)
        _raise(self._lib.vsplat_set_filter3d(self._h, _p(f), f.size), "set_filter3d")

    def set_sh_dropout(self, probability: float, retained_degree: int, seed: int):
        _raise(self._lib.vsplat_set_sh_dropout(self._h, probability, retained_degree, seed), "set_sh_dropout")

    def parameters(self, group: int):
        ptr, size = C.c_void_p(), C.c_uint64()
        _raise(self._lib.vsplat_parameters(self._h, group, C.byref(ptr), C.byref(size)), "parameters")
        return ptr.value, size.value

    # --- accessors
    def uses_fixed_point(self) -> bool:
        return bool(self._lib.vsplat_uses_fixed_point(self._h))

    def active_sh_degree(self) -> int:
        return int(self._lib.vsplat_active_sh_degree(self._h))

    def gaussian_count(self) -> int:
        return int(self._lib.vsplat_gaussian_count(self._h))

    def gpu_bytes(self) -> int:
        return int(self._lib.vsplat_gpu_bytes(self._h))

    def bytes_per_gaussian(self) -> int:
        return int(self._lib.vsplat_bytes_per_gaussian(self._h))

    def staging_bytes(self) -> int:
        return int(self._lib.vsplat_staging_bytes(self._h))

    def timings(self) -> list:
        items = (CTiming * 16)()
        count = C.c_uint32()
        _raise(self._lib.vsplat_timings(self._h, items, 16, C.byref(count)), "timings")
        return [(items[i].name.decode(), items[i].milliseconds) for i in range(min(count.value, 16))]

    def reset_timings(self):
        _raise(self._lib.vsplat_reset_timings(self._h), "reset_timings")

    def set_stage_timing(self, enabled: bool):
        _raise(self._lib.vsplat_set_stage_timing(self._h, int(bool(enabled))), "set_stage_timing")

    # --- the batched step (no reference counterpart)
    def step_views(self, cameras: list, targets: list, rates: LearningRates | None = None) -> list:
        """One iteration over several views: gradients summed in view order, then
        one Adam step when rates are given (else the sum stays for gradients())."""
        n = len(cameras)
        if n != len(targets):
            raise ValueError("one target per camera")
        cams = (CCamera * n)(*[camera_to_c(c) for c in cameras])
        arrays = [np.ascontiguousarray(t, np.uint8) for t in targets]
        ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in arrays])
        sizes = (C.c_size_t * n)(*[a.nbytes for a in arrays])
        stats = (CStats * n)()
        mode = MODE_STEP if rates is not None else MODE_GRADIENTS
        r = None if rates is None else rates.to_c()
        _raise(self._lib.vsplat_step_views(self._h, n, cams, ptrs, sizes, mode, None if r is None else C.byref(r), stats),
               "step_views")
        return [StepStats.of(s) for s in stats]

    def enqueue_views(self, cameras: list, device_target_ptrs: list, mode: int, rates: LearningRates | None = None):
        n = len(cameras)
        cams = (CCamera * n)(*[camera_to_c(c) for c in cameras])
        ptrs = (C.c_void_p * n)(*device_target_ptrs)
        r = None if rates is None else rates.to_c()
        _raise(self._lib.vsplat_enqueue_views(self._h, n, cams, ptrs, mode, None if r is None else C.byref(r)), "enqueue_views")

    def synchronize_views(self) -> list:
        stats = (CStats * 64)()
        count = C.c_uint32()
        _raise(self._lib.vsplat_synchronize_views(self._h, stats, 64, C.byref(count)), "synchronize_views")
        return [StepStats.of(stats[k]) for k in range(min(count.value, 64))]

    def capture_views(self, cameras: list, device_target_ptrs: list, mode: int, rates: LearningRates | None = None):
        """Captures the batch enqueue_views would queue as a HIP graph (not run)."""
        n = len(cameras)
        cams = (CCamera * n)(*[camera_to_c(c) for c in cameras])
        ptrs = (C.c_void_p * n)(*device_target_ptrs)
        r = None if rates is None else rates.to_c()
        _raise(self._lib.vsplat_capture_views(self._h, n, cams, ptrs, mode, None if r is None else C.byref(r)), "capture_views")

    def launch_captured(self, replays: int = 1):
        """Queues the captured batch `replays` times; wait with synchronize_views."""
        _raise(self._lib.vsplat_launch_captured(self._h, replays), "launch_captured")

    def autotune_tiles(self, camera: Camera, device_target_ptr: int, repeats: int = 2) -> tuple:
        """Times the tile shapes on this view, keeps the fastest for its size; returns (width, height)."""
        cam = camera_to_c(camera)
        w, h = C.c_uint32(), C.c_uint32()
        _raise(self._lib.vsplat_autotune_tiles(self._h, C.byref(cam), C.c_void_p(device_target_ptr), repeats, C.byref(w), C.byref(h)),
               "autotune_tiles")
        return w.value, h.value

    def autotune_views(self, cameras: list, device_target_ptrs: list, repeats: int = 2) -> list:
        """Tunes the tile shapes of the batch's view sizes together, timing the
        whole batch; returns each view's (width, height)."""
        n = len(cameras)
        cams = (CCamera * n)(*[camera_to_c(c) for c in cameras])
        ptrs = (C.c_void_p * n)(*device_target_ptrs)
        w, h = (C.c_uint32 * n)(), (C.c_uint32 * n)()
        _raise(self._lib.vsplat_autotune_views(self._h, n, cams, ptrs, repeats, w, h), "autotune_views")
        return [(w[k], h[k]) for k in range(n)]

    def release_captured(self):
        _raise(self._lib.vsplat_release_captured(self._h), "release_captured")

    # --- device-resident path
    def enqueue(self, camera: Camera, device_target_ptr: int, mode: int, rates: LearningRates | None = None):
        cam = camera_to_c(camera)
        r = None if rates is None else rates.to_c()
        _raise(self._lib.vsplat_enqueue(self._h, C.byref(cam), C.c_void_p(device_target_ptr), mode,
                                        None if r is None else C.byref(r)), "enqueue")

    def synchronize(self) -> StepStats:
        st = CStats()
        _raise(self._lib.vsplat_synchronize(self._h, C.byref(st)), "synchronize")
        return StepStats.of(st)

    def stream(self) -> int:
        return int(self._lib.vsplat_stream(self._h) or 0)

    def reserve_entries(self, entries: int):
        _raise(self._lib.vsplat_reserve_entries(self._h, entries), "reserve_entries")

    def inspect(self, name: str, view: int = 0) -> np.ndarray:
        """An intermediate buffer as the last call left it; `view` picks a batch's view."""
        size = C.c_uint64()
        L = self._lib
        _raise(L.vsplat_inspect_view(self._h, view, name.encode(), None, 0, C.byref(size)), "inspect")
        if name == "raw":
            dtype, width = (np.int64, 14) if self.uses_fixed_point() else (np.float32, 14)
        else:
            dtype, width = INSPECT_TYPES[name]
        out = np.zeros(size.value // np.dtype(dtype).itemsize, dtype)
        _raise(L.vsplat_inspect_view(self._h, view, name.encode(), _p(out), out.nbytes, C.byref(size)), "inspect")
        return out.reshape(-1, width) if width > 1 else out


def refined_pose(camera: Camera, move) -> Camera:
    """refined_pose (src/train/trainer.hpp:78-80) through the library: the camera
    moved by (rotation w, translation v) in camera space."""
    m = np.ascontiguousarray(move, np.float32).reshape(-1)
    if m.size != 6:
        raise ValueError("a move is 6 floats: w, then v")
    c, out = camera_to_c(camera), CCamera()
    _raise(lib().vsplat_refined_pose(C.byref(c), _p(m), C.byref(out)), "refined_pose")
    view = np.array(out.view[:], np.float32).reshape(4, 4).T.copy()
    return Camera(view, np.array(out.position[:], np.float32), out.fx, out.fy, out.cx, out.cy, out.width, out.height)


def load_ply(path) -> Cloud:
    """load_ply (src/splat/gaussian_cloud.hpp:60-63) through the library."""
    p = str(path).encode()
    count, degree = C.c_uint64(), C.c_int32()
    _raise(lib().vsplat_ply_info(p, C.byref(count), C.byref(degree)), "load_ply")
    c = Cloud.zeros(count.value, degree.value)
    _raise(lib().vsplat_read_ply(p, _p(c.positions), _p(c.scales), _p(c.rotations), _p(c.opacities), _p(c.sh), count.value),
           "load_ply")
    return c


def save_ply(cloud: Cloud, path) -> None:
    """save_ply (src/splat/gaussian_cloud.hpp:65-67) through the library."""
    arrays = [np.ascontiguousarray(a, np.float32) for a in
              (cloud.positions, cloud.scales, cloud.rotations, cloud.opacities, cloud.sh)]
    cc = CCloud(*[_p(a) for a in arrays], cloud.count, cloud.sh_degree)
    _raise(lib().vsplat_write_ply(str(path).encode(), C.byref(cc)), "save_ply")


class DeviceBuffer:
    """GPU memory holding a copy of a host array, for the device-resident calls
    (Trainer.enqueue takes its .ptr as the target). Allocated through the
    library, so that it shares the library's HIP runtime."""

    def __init__(self, array: np.ndarray):
        a = np.ascontiguousarray(array)
        ptr = C.c_void_p()
        _raise(lib().vsplat_device_alloc(a.nbytes, C.c_void_p(a.ctypes.data), C.byref(ptr)), "device_alloc")
        self.ptr, self.nbytes = ptr.value, a.nbytes

    @classmethod
    def empty(cls, nbytes: int) -> "DeviceBuffer":
        """Uninitialised GPU memory of nbytes."""
        self = cls.__new__(cls)
        ptr = C.c_void_p()
        _raise(lib().vsplat_device_alloc(nbytes, None, C.byref(ptr)), "device_alloc")
        self.ptr, self.nbytes = ptr.value, nbytes
        return self

    def read(self, dtype=np.uint8) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        _raise(lib().vsplat_device_read(C.c_void_p(self.ptr), _p(out), out.nbytes), "device_read")
        return out

    def copy_from(self, other: "DeviceBuffer", repeats: int = 1) -> float:
        """Copies `other` into this buffer `repeats` times; returns ms per copy."""
        if other.nbytes != self.nbytes:
            raise ValueError("copy between buffers of different sizes")
        ms = C.c_float()
        _raise(lib().vsplat_device_copy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), self.nbytes, repeats, C.byref(ms)),
               "device_copy")
        return ms.value

    def free(self):
        if getattr(self, "ptr", None):
            lib().vsplat_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001 (interpreter shutdown)
            pass


def math_probe(function: int, x: np.ndarray) -> np.ndarray:
    """The device's exp (0), log (1), sqrt (2), reciprocal (3), half round trip (4), half bits to float (5) or
    the backward's division of pairs (6: out[2k] = x[2k] / x[2k + 1] in its stated range, out[2k + 1] = 0)."""
    a = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(a)
    _raise(lib().vsplat_math_probe(function, _p(a), _p(out), a.size), "math_probe")
    return out
