#!/usr/bin/env python3
"""fwd+bwd ms/iter of the differentiable tile rasterizer on one MI355X.

Workload (the 1M-Gaussian, 1280x720 configuration the metric is quoted on):
1,000,000 Gaussians, SH degree 3 (48 coefficients N(0, 0.3)), means U[-1,1]^3,
log-scales N(-4, 0.5), normalised N(0,1) quaternions, opacity logit N(0,1); one
1280x720 camera at (0,0,-3) looking at the origin, 60 degree vertical field of
view; loss 0.8 L1 + 0.2 D-SSIM against a smooth synthetic target; 64-bit
fixed-point backward. Inputs are generated from seeded explicit RNG maths
(vsplat.fixtures) and are resident in GPU memory before the timed region.

A step is one pass of the hot path: projection and SH, depth sort, tile
binning and the tile sort, compositing, L1+SSIM and its adjoint, the backward
pass and the chain rule to fp32 gradients in GPU memory, queued on the
trainer's stream with no host synchronisation inside (vsplat_enqueue).
W untimed steps, then K timed ones bracketed by synchronisation; per-stage
times come from HIP events recorded on the trainer's stream.

Prints one JSON line. --gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "vibe-splat_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import vsplat  # noqa: E402
from vsplat import fixtures as F  # noqa: E402

METRIC = "fwd+bwd ms/iter, 1M Gaussians @1280x720; batched-bf16 vs plain fp32; % HBM peak"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def touched_entries(ranges: np.ndarray, stops: np.ndarray, width: int, height: int, tiles=(16, 16)) -> int:
    """Ep = sum over tiles of (range end - the lowest pixel stop): the entries the
    forward and backward passes must visit."""
    tw, th = tiles
    gw, gh = (width + tw - 1) // tw, (height + th - 1) // th
    s = stops.reshape(height, width)
    pad = np.full((gh * th, gw * tw), np.iinfo(np.uint32).max, np.uint64)
    pad[:height, :width] = s
    low = pad.reshape(gh, th, gw, tw).min(axis=(1, 3)).reshape(-1)
    end = ranges[:, 1].astype(np.int64)
    low = np.minimum(low.astype(np.int64), end)
    return int((end - low).sum())


def algorithmic_bytes(N, V, E, Ep, P, T, Bp):
    """Compulsory HBM traffic of one iteration: each stage reads its logical
    inputs once and writes its outputs once (sort passes, histograms, LDS
    staging and halo re-reads not counted)."""
    stages = {
        "project": N * Bp + 56 * V,
        "depth sort": 16 * V,
        "binning": 56 * V + 8 * E,
        "tile sort": 16 * E,
        "ranges": 4 * E + 8 * T,
        "forward": 52 * Ep + 8 * T + 20 * P,
        "loss": 132 * P,
        "backward": 52 * Ep + 8 * T + 24 * P + 56 * V,
        "chain rule": 56 * V + N * Bp + 236 * N,
    }
    return sum(stages.values()), stages


def cpu_baseline(workload, threads: int) -> dict:
    """The reference's CPU path (its src/train/reference.cpp, as the FP64 oracle
    restates it): forward, loss and backward with the chain rule over the whole
    1280x720 view, timed once on `threads` host threads. Beside it: the same on
    one thread, from a 12-row and a 2-row band of the view (the per-row work
    scaled to 720 rows: a whole single-thread iteration takes about a minute), and
    the FP32 model of this build's GPU algorithm on one thread, a whole iteration."""
    from oracle import oracle as O

    cam, target = workload.cameras[0], workload.targets[0]
    H = cam.height

    def band(rows):
        top = (H - rows) // 2
        return top, dataclasses.replace(cam, height=rows, cy=float(np.float32(cam.cy) - np.float32(top)))

    def oracle_a(rows, n):
        top, c = band(rows)
        t0 = time.perf_counter()
        img = O.render(workload.cloud, c, threads=n)
        ls = O.loss(img, np.ascontiguousarray(target[top:top + rows]), workload.lambda_ssim, threads=n)
        O.gradients(workload.cloud, c, ls.adjoint, threads=n)
        return time.perf_counter() - t0

    def oracle_b():
        t0 = time.perf_counter()
        m = O.Model(workload.cloud, cam)
        r = m.forward()
        m.loss(r, target, workload.lambda_ssim)
        m.backward_fixed(r)
        m.gradients(r)
        return time.perf_counter() - t0

    full = oracle_a(H, threads)
    one_fixed, one_band = oracle_a(2, 1), oracle_a(12, 1)
    one = one_fixed + (one_band - one_fixed) / 10.0 * (H - 2)
    b_full = oracle_b()
    return {"value": round(full * 1000.0, 1), "unit": "ms/iter", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(),
            "sample": f"FP64 restatement of reference.cpp, one whole iteration (fwd+loss+bwd+chain rule) of the "
                      f"1280x720 view on {threads} threads",
            "single_core": {"value": round(one * 1000.0, 1), "unit": "ms/iter", "cores": 1,
                            "sample": f"the same on 1 thread, from a 12-row band ({one_band:.2f} s) and a 2-row band "
                                      f"({one_fixed:.2f} s) of the view, the per-row time scaled to 720 rows"},
            "fp32_tile_model": {"value": round(b_full * 1000.0, 1), "unit": "ms/iter", "cores": 1,
                                "sample": "the FP32 model of this build's GPU algorithm (same lists, sums and fixed "
                                          "point), one whole iteration on 1 thread"}}


def fp64_errors(forms, crop_size: int = 128) -> dict:
    """Each timed form's errors against the FP64 restatement of the reference
    (the oracle as the checker), on a crop_size x crop_size crop at the centre
    of its view: the image, the loss, and each parameter group's gradient as a
    relative L2 error. The GPU runs the form's own options (its SH storage and
    its fixed-point sums); the FP64 side takes the cloud as the trainer stores it."""
    from oracle import oracle as O

    out = {}
    for name, cloud, cam, options in forms:
        crop = dataclasses.replace(cam, width=crop_size, height=crop_size, cx=cam.cx - (cam.width - crop_size) / 2,
                                   cy=cam.cy - (cam.height - crop_size) / 2)
        target = F.smooth_target_explicit(crop_size, crop_size, 1234)
        stored = cloud.copy()
        stored.sh = F.unpack_sh(F.pack_sh(cloud.sh, options.sh_storage), cloud.coefficients,
                                options.sh_storage).reshape(cloud.sh.shape).copy()
        ref = O.render(stored, crop)
        ref_loss = O.loss(ref, target, options.lambda_ssim)
        ref_g = O.gradients(stored, crop, ref_loss.adjoint)
        with vsplat.Trainer(cloud, options) as t:
            image = t.render(crop)
            st = t.compute_gradients(crop, target)
            g = t.gradients()
        errors = {}
        for group in ("positions", "scales", "rotations", "opacities", "sh"):
            a = np.asarray(getattr(g, group), np.float64).ravel()
            b = np.asarray(getattr(ref_g, group), np.float64).ravel()
            errors[group] = float(f"{np.linalg.norm(a - b) / np.linalg.norm(b):.3e}")
        out[name] = {"crop": f"{crop_size}x{crop_size} at the centre of {cam.width}x{cam.height}",
                     "image_max_abs": float(f"{np.abs(image[:, :3] - ref.rgb).max():.3e}"),
                     "loss_abs": float(f"{abs(st.loss - ref_loss.loss):.3e}"),
                     "gradient_rel_l2": errors, "gradient_rel_l2_max": max(errors.values())}
    return out


def apply_changes_leg(trainer, n: int) -> dict:
    """Trainer::apply_changes after a densification round on the benchmark
    trainer (run last: it changes the cloud): every tenth Gaussian pruned and
    every tenth copied, first with the values gathered on the device (no
    transfer), then the same table with the values handed over from the host
    as densify() does (the cloud downloaded, the new one uploaded). Wall time of
    the synchronous calls, allocation of the new arrays included."""
    index = np.arange(n)
    source = np.concatenate([index[index % 10 != 3], index[::10]]).astype(np.int64)
    t0 = time.perf_counter()
    trainer.apply_changes(None, source)
    device_ms = (time.perf_counter() - t0) * 1000.0
    m = len(source)
    t0 = time.perf_counter()
    cloud = trainer.cloud()
    trainer.apply_changes(cloud, np.arange(m))
    host_ms = (time.perf_counter() - t0) * 1000.0
    # Per new Gaussian: parameters as stored (148 bytes at SH 3) and Adam's state
    # (4 bytes per parameter, 236), each read once and written once.
    return {"gaussians_before": n, "gaussians_after": m, "device_ms": round(device_ms, 3),
            "host_cloud_ms": round(host_ms, 3), "algorithmic_bytes_device": 2 * m * (148 + 236)}


def ply_leg(cloud) -> dict:
    """3DGS .ply writing and reading of the benchmark cloud on the host
    (vsplat_write_ply, vsplat_read_ply), through a temporary file."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cloud.ply")
        t0 = time.perf_counter()
        vsplat.save_ply(cloud, path)
        write_ms = (time.perf_counter() - t0) * 1000.0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        loaded = vsplat.load_ply(path)
        read_ms = (time.perf_counter() - t0) * 1000.0
    ok = all(np.array_equal(getattr(loaded, k), getattr(cloud, k)) for k in ("positions", "scales", "rotations", "opacities", "sh"))
    return {"gaussians": cloud.count, "bytes": size, "write_ms": round(write_ms, 1), "read_ms": round(read_ms, 1),
            "write_gbs": round(size / (write_ms * 1e-3) / 1e9, 3), "read_gbs": round(size / (read_ms * 1e-3) / 1e9, 3),
            "round_trip_exact": bool(ok)}


def copy_rate(nbytes: int = 1 << 30, repeats: int = 20) -> dict:
    """The device's measured copy rate: one GiB copied between two GPU buffers,
    back to back, timed with HIP events (read + write bytes per second)."""
    src, dst = vsplat.DeviceBuffer.empty(nbytes), vsplat.DeviceBuffer.empty(nbytes)
    try:
        ms = dst.copy_from(src, repeats)
    finally:
        src.free()
        dst.free()
    return {"bytes": nbytes, "ms_per_copy": round(ms, 4), "gbs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


# Timed passes per tile shape when tuning (the least kept; the tuning is not
# timed): shapes within a few per cent of each other need several to be told apart.
TUNE_REPEATS = 6
TILE_SHAPES = {"auto": vsplat.TILES_AUTO, "16x16": vsplat.TILES_16X16, "8x8": vsplat.TILES_8X8, "32x8": vsplat.TILES_32X8,
               "32x32": vsplat.TILES_32X32}


def batched_ring(steps: int, warmup: int, count: int | None, tile_shape: str) -> dict:
    """The batched form on the eight-view ring (1M Gaussians, SH 3 stored as
    bf16, views of 640x480 to 1920x1080 on a ring of radius 3), fixed point, with Adam: each
    view timed as its own step (vsplat_enqueue), then the eight as one batched
    iteration (vsplat_enqueue_views: gradients summed over the views, one Adam
    step). Targets are in GPU memory; no stage events in either timed loop."""
    w = F.config(3, count=count)
    rates = vsplat.LearningRates()
    trainer = vsplat.Trainer(w.cloud, vsplat.TrainerOptions(lambda_ssim=w.lambda_ssim, force_fixed_point=True,
                                                            sh_storage=vsplat.SH_BF16, tile_shape=TILE_SHAPES[tile_shape]))
    trainer.set_stage_timing(False)
    buffers = [vsplat.DeviceBuffer(t) for t in w.targets]
    ptrs = [b.ptr for b in buffers]
    shapes = {}
    try:
        if tile_shape == "auto":  # the sizes' shapes tuned together, on the batch
            for cam, shape in zip(w.cameras, trainer.autotune_views(w.cameras, ptrs, TUNE_REPEATS)):
                shapes[(cam.width, cam.height)] = shape
        trainer.step_views(w.cameras, w.targets, rates)  # sizes every buffer and the entry reservation
        single = []
        for cam, target, ptr in zip(w.cameras, w.targets, ptrs):
            trainer.step(cam, target, rates)  # sizes the single-view buffers for this view
            for _ in range(warmup):
                trainer.enqueue(cam, ptr, vsplat.MODE_STEP, rates)
            trainer.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                trainer.enqueue(cam, ptr, vsplat.MODE_STEP, rates)
            trainer.synchronize()
            single.append((time.perf_counter() - t0) * 1000.0 / steps)
        for _ in range(warmup):
            trainer.enqueue_views(w.cameras, ptrs, vsplat.MODE_STEP, rates)
            trainer.synchronize_views()
        t0 = time.perf_counter()
        for _ in range(steps):
            trainer.enqueue_views(w.cameras, ptrs, vsplat.MODE_STEP, rates)
        trainer.synchronize_views()
        batch = (time.perf_counter() - t0) * 1000.0 / steps
        # The same iteration captured once as a HIP graph and replayed.
        trainer.capture_views(w.cameras, ptrs, vsplat.MODE_STEP, rates)
        trainer.launch_captured(warmup)
        trainer.synchronize_views()
        t0 = time.perf_counter()
        trainer.launch_captured(steps)
        trainer.synchronize_views()
        graph = (time.perf_counter() - t0) * 1000.0 / steps
        trainer.release_captured()
    finally:
        for b in buffers:
            b.free()
        trainer.close()
    views = len(w.cameras)
    return {"workload": w.description, "views": views, "sizes": [[c.width, c.height] for c in w.cameras],
            "single_view_ms": [round(x, 4) for x in single], "single_views_total_ms": round(sum(single), 4),
            "batched_ms": round(batch, 4), "batched_ms_per_view": round(batch / views, 4),
            "speedup_per_view": round(sum(single) / batch, 4),
            "graph_ms": round(graph, 4), "graph_ms_per_view": round(graph / views, 4),
            "graph_speedup_per_view": round(sum(single) / graph, 4), "sh_storage": "bf16 (all 48 coefficients), fp32 arithmetic",
            "tile_shape": tile_shape, "tiles_chosen": {f"{k[0]}x{k[1]}": f"{v[0]}x{v[1]}" for k, v in shapes.items()}}


def ring_1280x720(views: int):
    """Eight (views) 1280x720 cameras on a ring of radius 3 about the cloud,
    the first the single view's camera, and their targets."""
    import math
    cams = [F.look_at((3.0 * math.sin(2.0 * math.pi * k / views), 0.0, -3.0 * math.cos(2.0 * math.pi * k / views)), (0, 0, 0),
                      (0.0, -1.0, 0.0), 60.0, 1280, 720) for k in range(views)]
    targets = [F.smooth_target_explicit(1280, 720, 1016 + k) for k in range(views)]
    return cams, targets


def plain_ring_1280x720(steps: int, warmup: int, count: int | None, tile_shape: str, fixed: bool, views: int = 8) -> dict:
    """The plain form (fp32 SH, one view per pass, no graph) on each camera of
    the batched 1280x720 ring in turn: ms per iteration of each view, gradients
    mode, and their mean, the like-for-like baseline of the batched form."""
    w = F.config(1, count=count)
    cams, targets = ring_1280x720(views)
    trainer = vsplat.Trainer(w.cloud, vsplat.TrainerOptions(lambda_ssim=w.lambda_ssim, force_fixed_point=fixed,
                                                            tile_shape=TILE_SHAPES[tile_shape]))
    trainer.set_stage_timing(False)
    per_view = []
    try:
        for cam, target in zip(cams, targets):
            buffer = vsplat.DeviceBuffer(target)
            try:
                if tile_shape == "auto":
                    trainer.autotune_tiles(cam, buffer.ptr, TUNE_REPEATS)
                trainer.compute_gradients(cam, target)
                for _ in range(warmup):
                    trainer.enqueue(cam, buffer.ptr, vsplat.MODE_GRADIENTS)
                trainer.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    trainer.enqueue(cam, buffer.ptr, vsplat.MODE_GRADIENTS)
                trainer.synchronize()
                per_view.append(round((time.perf_counter() - t0) * 1000.0 / steps, 4))
            finally:
                buffer.free()
    finally:
        trainer.close()
    return {"ms_per_view": per_view, "mean_ms_per_view": round(sum(per_view) / len(per_view), 4)}


def batched_1280x720(steps: int, warmup: int, count: int | None, tile_shape: str, views: int = 8, fixed: bool = True,
                     schedule: int = vsplat.BATCH_VIEW_STREAMS) -> dict:
    """The target comparison at one size: the 1M-Gaussian cloud of `value`, seen
    by eight 1280x720 views on a ring of radius 3 (the ring workload's cameras,
    all at 1280x720), gradients mode (fwd+bwd, gradients summed over the views),
    SH stored as bf16, the sums in 64-bit fixed point (fixed) or float32
    atomics, captured as one HIP graph and replayed; per view against the plain
    fp32 single 1280x720 view."""
    w = F.config(1, count=count)
    cams, targets = ring_1280x720(views)
    buffers = [vsplat.DeviceBuffer(t) for t in targets]
    ptrs = [b.ptr for b in buffers]

    def graph_ms(shape: str, replays: int, before: int) -> float:
        trainer = vsplat.Trainer(w.cloud, vsplat.TrainerOptions(lambda_ssim=w.lambda_ssim, force_fixed_point=fixed,
                                                                sh_storage=vsplat.SH_BF16, tile_shape=TILE_SHAPES[shape],
                                                                batch_raster=schedule))
        trainer.set_stage_timing(False)
        try:
            trainer.step_views(cams, targets)  # gradients mode; sizes every view's buffers
            trainer.capture_views(cams, ptrs, vsplat.MODE_GRADIENTS)
            trainer.launch_captured(before)
            trainer.synchronize_views()
            t0 = time.perf_counter()
            trainer.launch_captured(replays)
            trainer.synchronize_views()
            ms = (time.perf_counter() - t0) * 1000.0 / replays
            trainer.release_captured()
        finally:
            trainer.close()
        return ms

    chosen = None
    try:
        shape = tile_shape
        if tile_shape == "auto":
            # The batch is measured as a captured graph, so its tile shape is chosen
            # by replaying each shape's graph (not timed). The library's tuner times
            # eager passes, in which 16x16 and 32x32 come out within noise of each
            # other at this size while 32x32's graph replays about 3 % faster.
            tuned = {s: graph_ms(s, 5, 2) for s in ("16x16", "8x8", "32x8", "32x32")}
            shape = chosen = min(tuned, key=tuned.get)
        graph = graph_ms(shape, steps, warmup)
    finally:
        for b in buffers:
            b.free()
    return {"workload": f"1M Gaussians, SH degree 3 stored as bf16, {views} views of 1280x720 on a ring of radius 3, "
                        "gradients mode, one captured graph", "sums": "64-bit fixed point" if fixed else "float32 atomics",
            "schedule": "one forward and one backward launch over all views" if schedule == vsplat.BATCH_ONE_LAUNCH
            else "each view's passes on its own stream",
            "views": views, "graph_ms": round(graph, 4),
            "graph_ms_per_view": round(graph / views, 4), "tile_shape": tile_shape,
            "tiles_chosen": chosen}


def other_form(index: int, steps: int, warmup: int, tile_shape: str, fixed: bool = True, **options) -> dict:
    """ms/iter of one more single-view workload, gradients mode, fixed point,
    target in device memory: the precision form (3M Gaussians in 64 blobs, SH
    stored as bf16, 1920x1080) or the differentiation stress (6M heavy-tailed
    Gaussians, 1920x1080, the backward recomputing the transmittance)."""
    w = F.config(index)
    cam, target = w.cameras[0], w.targets[0]
    trainer = vsplat.Trainer(w.cloud, vsplat.TrainerOptions(lambda_ssim=w.lambda_ssim, force_fixed_point=fixed,
                                                            tile_shape=TILE_SHAPES[tile_shape], **options))
    trainer.set_stage_timing(False)
    buffer = vsplat.DeviceBuffer(target)
    chosen = None
    try:
        if tile_shape == "auto":
            chosen = trainer.autotune_tiles(cam, buffer.ptr, TUNE_REPEATS)
        first = trainer.compute_gradients(cam, target)
        for _ in range(warmup):
            trainer.enqueue(cam, buffer.ptr, vsplat.MODE_GRADIENTS)
        trainer.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            trainer.enqueue(cam, buffer.ptr, vsplat.MODE_GRADIENTS)
        st = trainer.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        trainer.set_stage_timing(True)
        trainer.reset_timings()
        for _ in range(max(3, steps // 4)):
            trainer.enqueue(cam, buffer.ptr, vsplat.MODE_GRADIENTS)
        trainer.synchronize()
        stages = {k: round(v, 4) for k, v in trainer.timings()}
        gpu_gb = trainer.gpu_bytes() / 1e9
    finally:
        buffer.free()
        trainer.close()
    return {"workload": w.description, "gaussians": w.cloud.count, "width": cam.width, "height": cam.height,
            "options": {k: (int(v) if isinstance(v, bool) else v) for k, v in options.items()},
            "sums": "64-bit fixed point" if fixed else "float32 atomics", "ms_per_iter": round(ms, 4),
            "visible": st.visible_splats, "entries": st.tile_entries, "loss": round(float(first.loss), 6),
            "gpu_gb": round(gpu_gb, 3), "tile_shape": tile_shape,
            "tiles_chosen": None if chosen is None else f"{chosen[0]}x{chosen[1]}", "stages_ms": stages}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--count", type=int, default=None, help="Gaussians (default: the workload's 1,000,000)")
    ap.add_argument("--mode", choices=["gradients", "step"], default="gradients")
    ap.add_argument("--float-atomics", action="store_true", help="float-atomic backward instead of fixed point")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-batched", action="store_true", help="skip the eight-view batched measurement")
    ap.add_argument("--no-other-forms", action="store_true", help="skip the 3M bf16 and 6M recompute workloads")
    ap.add_argument("--no-densify", action="store_true", help="skip the apply_changes measurement")
    ap.add_argument("--tile-shape", choices=sorted(TILE_SHAPES), default="auto",
                    help="raster tiles (auto: timed per view size; every shape gives the same results)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU view split is a design note, not code)", file=sys.stderr)
        sys.exit(2)

    log("generating the workload")
    w = F.config(1, count=args.count)
    cam, target = w.cameras[0], w.targets[0]
    N, P = w.cloud.count, cam.width * cam.height
    fixed = not args.float_atomics
    options = vsplat.TrainerOptions(lambda_ssim=w.lambda_ssim, force_fixed_point=fixed, tile_shape=TILE_SHAPES[args.tile_shape])
    trainer = vsplat.Trainer(w.cloud, options)
    device_target = vsplat.DeviceBuffer(target)
    d_target = device_target.ptr
    if args.tile_shape == "auto":
        log("tile shape: timing 16x16, 8x8, 32x8 and 32x32")
        tiles = trainer.autotune_tiles(cam, d_target, TUNE_REPEATS)
        log(f"  chose {tiles[0]}x{tiles[1]}")
    else:
        tiles = tuple(int(v) for v in args.tile_shape.split("x"))
    rates = vsplat.LearningRates()
    mode = vsplat.MODE_GRADIENTS if args.mode == "gradients" else vsplat.MODE_STEP

    log("sizing: one synchronous step")
    first = trainer.compute_gradients(cam, target)
    log(f"  loss {first.loss:.6f} visible {first.visible_splats} entries {first.tile_entries}")
    for _ in range(args.warmup):
        trainer.enqueue(cam, d_target, mode, rates if mode == vsplat.MODE_STEP else None)
    trainer.synchronize()
    trainer.reset_timings()

    def timed(stage_events: bool):
        trainer.set_stage_timing(stage_events)
        trainer.reset_timings()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            trainer.enqueue(cam, d_target, mode, rates if mode == vsplat.MODE_STEP else None)
        st = trainer.synchronize()
        return st, (time.perf_counter() - t0) * 1000.0 / args.steps

    log(f"timing {args.steps} steps")
    stats, ms = timed(False)
    log(f"  {ms:.4f} ms per step")
    log(f"timing {args.steps} steps with per-stage HIP events")
    _, ms_staged = timed(True)
    stages = dict(trainer.timings())
    log(f"  {ms_staged:.4f} ms per step: " + ", ".join(f"{k} {v:.3f}" for k, v in stages.items()))

    V, E = stats.visible_splats, stats.tile_entries
    T = ((cam.width + tiles[0] - 1) // tiles[0]) * ((cam.height + tiles[1] - 1) // tiles[1])
    Ep = touched_entries(trainer.inspect("ranges")[:T], trainer.inspect("stops"), cam.width, cam.height, tiles)
    coefficients = (w.cloud.sh_degree + 1) ** 2
    Bp = (3 + 3 + 4 + 1) * 4 + F.half_sh_words(coefficients) * 4  # as stored: SH DC fp32, rest fp16
    total, per_stage = algorithmic_bytes(N, V, E, Ep, P, T, Bp)
    dominant = max(("composite", "forward"), ("backward", "backward"), key=lambda s: stages.get(s[0], 0.0))
    kernel_ms = stages[dominant[0]]
    achieved = per_stage[dominant[1]] / (kernel_ms * 1e-3) / 1e9
    log("device copy rate")
    copy = copy_rate()
    log(f"  {copy['gbs']} GB/s")
    # HBM bytes per launch of that kernel, from the FETCH_SIZE / WRITE_SIZE counter runs
    # on this workload (profiles/), when this run is that workload.
    # The same runs' SQ counters give the kernel's share of the chip's VALU issue
    # slots and of its LDS bandwidth.
    traffic, issue = None, {}
    traffic_file = ROOT / "profiles" / "traffic_1M_1280x720_fixed_gradients.json"
    kernel_name = {"forward": f"forward_kernel<{tiles[0]}u, {tiles[1]}u>",
                   "backward": f"backward_kernel<true, false, false, {tiles[0]}u, {tiles[1]}u>"}[dominant[1]]
    if traffic_file.exists() and fixed and N == 1_000_000 and args.mode == "gradients":
        entry = json.loads(traffic_file.read_text())["kernels"].get(kernel_name)
        if entry is not None:
            traffic = entry["hbm_bytes_per_launch"]
            issue = {"valu_busy": entry.get("valu_busy"), "lds_frac": entry.get("lds_frac"),
                     "counters": f"profiles/{traffic_file.name}"}

    result = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "1M Gaussians SH3, one 1280x720 view, 0.8 L1 + 0.2 D-SSIM, "
                               + ("fixed-point" if fixed else "float-atomic") + " backward, " + args.mode,
                   "gaussians": N, "width": cam.width, "height": cam.height, "sh_degree": w.cloud.sh_degree,
                   "form": "plain fp32, single view", "tile_shape": args.tile_shape, "tiles_chosen": f"{tiles[0]}x{tiles[1]}"},
        "roofline": {"bound": "hbm", "kernel": dominant[1], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "algorithmic_bytes": int(per_stage[dominant[1]]), "kernel_ms": round(kernel_ms, 4),
                     "measured_copy_gbs": copy["gbs"], "frac_of_measured_copy": round(achieved / copy["gbs"], 5), **issue},
        "iteration": {"algorithmic_bytes": int(total), "hbm_frac": round(total / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                      "frac_of_measured_copy": round(total / (ms * 1e-3) / 1e9 / copy["gbs"], 5),
                      "V": V, "E": E, "Ep": Ep, "P": P, "T": T, "Bp": Bp, "loss": round(float(stats.loss), 6)},
        "ms_per_step_with_stage_events": round(ms_staged, 4),
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
    }
    if not args.no_batched:
        log("batched form: the eight-view ring")
        result["batched"] = batched_ring(max(5, args.steps // 2), 2, args.count, args.tile_shape)
        bt = result["batched"]
        # The headline ratio: the plain fp32 single 1280x720 view against the
        # batched bf16 form's graph, per view (the ring's views average 1.25x its pixels).
        bt["plain_single_view_ms"] = result["value"]
        bt["plain_vs_graph_per_view"] = round(result["value"] / bt["graph_ms_per_view"], 4)
        log(f"  per view {bt['batched_ms_per_view']} ms batched (speedup {bt['speedup_per_view']}), "
            f"{bt['graph_ms_per_view']} ms as a graph (speedup {bt['graph_speedup_per_view']})")
        log("plain fp32 fwd+bwd on each camera of the eight-view 1280x720 ring in turn")
        plain_ring = plain_ring_1280x720(max(5, args.steps // 2), 2, args.count, args.tile_shape, fixed)
        result["plain_ring_1280x720"] = plain_ring
        log(f"  {plain_ring['ms_per_view']} ms, mean {plain_ring['mean_ms_per_view']} ms per view")
        log("the same with the sums as float atomics")
        plain_ring_float = plain_ring_1280x720(max(5, args.steps // 2), 2, args.count, args.tile_shape, False)
        result["plain_ring_1280x720_float_sums"] = plain_ring_float
        log(f"  {plain_ring_float['ms_per_view']} ms, mean {plain_ring_float['mean_ms_per_view']} ms per view")

        def compare(leg, float_sums=False):
            # Per view against the plain single view of `value`, and against the
            # plain form on the same eight cameras with the same kind of sums.
            same = plain_ring_float if float_sums else plain_ring
            leg["plain_single_view_ms"] = result["value"]
            leg["plain_vs_batched_per_view"] = round(result["value"] / leg["graph_ms_per_view"], 4)
            leg["plain_same_views_ms_per_view"] = same["mean_ms_per_view"]
            leg["plain_same_views_vs_batched_per_view"] = round(same["mean_ms_per_view"] / leg["graph_ms_per_view"], 4)
            return (f"{leg['graph_ms_per_view']} ms per view, {leg['plain_vs_batched_per_view']}x the plain single view, "
                    f"{leg['plain_same_views_vs_batched_per_view']}x the plain form on the same views with the same sums")

        log("batched bf16 fwd+bwd: eight 1280x720 views")
        same = batched_1280x720(max(5, args.steps // 2), 2, args.count, args.tile_shape)
        result["batched_1280x720"] = same
        log("  " + compare(same))
        log("batched bf16 fwd+bwd, sums accumulated in float32: eight 1280x720 views")
        same_float = batched_1280x720(max(5, args.steps // 2), 2, args.count, args.tile_shape, fixed=False)
        result["batched_1280x720_float_sums"] = same_float
        log("  " + compare(same_float, True))
        log("batched bf16 fwd+bwd, one forward and one backward launch over all views: eight 1280x720 views")
        for key, fixed_sums in (("batched_1280x720_one_launch", True), ("batched_1280x720_one_launch_float_sums", False)):
            one = batched_1280x720(max(5, args.steps // 2), 2, args.count, args.tile_shape, fixed=fixed_sums,
                                   schedule=vsplat.BATCH_ONE_LAUNCH)
            result[key] = one
            log(f"  {key}: " + compare(one, not fixed_sums))
    if not args.no_other_forms and args.count is None:
        result["other_forms"] = {}
        for key, index, fixed_sums, form_options in (
                ("precision_bf16_3M_1920x1080", 2, True, {"sh_storage": vsplat.SH_BF16}),
                ("precision_bf16_3M_1920x1080_float_sums", 2, False, {"sh_storage": vsplat.SH_BF16}),
                ("stress_recompute_6M_1920x1080", 4, True, {"recompute_transmittance": True})):
            log(f"{key}: generating and timing")
            form = other_form(index, max(5, args.steps // 2), 2, args.tile_shape, fixed=fixed_sums, **form_options)
            result["other_forms"][key] = form
            log(f"  {form['ms_per_iter']} ms per iteration, {form['entries']} entries, {form['gpu_gb']} GB")
    if not args.no_densify:
        log("ply: write and read the cloud")
        result["ply"] = ply_leg(w.cloud)
        log(f"  write {result['ply']['write_ms']} ms, read {result['ply']['read_ms']} ms for {result['ply']['bytes']} bytes")
        log("apply_changes: 10 % pruned, 10 % copied")
        result["apply_changes"] = apply_changes_leg(trainer, N)
        log(f"  {result['apply_changes']['device_ms']} ms on the device, "
            f"{result['apply_changes']['host_cloud_ms']} ms with the values from the host")
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"CPU baseline on {threads} threads")
        result["cpu_baseline"] = cpu_baseline(w, threads)
        # The same CPU path checks the timed forms: their errors against FP64.
        log("errors of the timed forms against FP64 (128x128 crops)")
        bf16 = dataclasses.replace(options, sh_storage=vsplat.SH_BF16)
        bf16_float = dataclasses.replace(bf16, force_fixed_point=False)
        forms = [("plain_fp32_fixed_1M_1280x720", w.cloud, cam, options),
                 ("batched_bf16_fixed_1M_1280x720", w.cloud, cam, bf16),
                 ("batched_bf16_float_sums_1M_1280x720", w.cloud, cam, bf16_float)]
        if args.count is None and not args.no_other_forms:
            w2 = F.config(2)
            for name, o in (("precision_bf16_fixed_3M_1920x1080", bf16), ("precision_bf16_float_sums_3M_1920x1080", bf16_float)):
                forms.append((name, w2.cloud, w2.cameras[0], dataclasses.replace(o, lambda_ssim=w2.lambda_ssim)))
        result["fp64_errors"] = fp64_errors(forms)
        for k, v in result["fp64_errors"].items():
            log(f"  {k}: gradient rel L2 up to {v['gradient_rel_l2_max']:.2e}, image {v['image_max_abs']:.1e}")
    device_target.free()
    trainer.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
