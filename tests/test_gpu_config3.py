"""The batching form at its own workload: 8 views of 640x480 to 1920x1080 on a
ring of radius 3, SH degree 3 stored as bf16, tile shapes chosen by the batch
autotuner, one fwd+bwd+Adam step captured as a HIP graph.

* At 20k Gaussians, bit for bit against Oracle B: each view's fixed-point sums
  and chain rule, the gradients added in view order in float32, then Oracle
  B's Adam (ob_adam) from the sum; each view's visible count and entries;
  under both batch schedules (the views' passes on their own streams, or one
  forward and one backward launch over all views' 8x8 blocks per tile shape).
* At the full 1M Gaussians, through properties: each view's tile lists
  sorted by tile and, within a tile, in depth order (the stable sort), the
  ranges partitioning the entries; two trainers replaying the captured step
  twice end bit for bit equal; each view's counts equal a single-view pass's.
"""
import dataclasses
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

RATES = vsplat.LearningRates(position=1.6e-4, scale=0.005, rotation=0.001, opacity=0.05, sh=0.025, sh_rest_scale=0.05)
RATE_TUPLE = (RATES.position, RATES.scale, RATES.rotation, RATES.opacity, RATES.sh)
SIZES = [(640, 480), (960, 540), (1280, 720), (1280, 720), (1280, 720), (1600, 900), (1920, 1080), (1920, 1080)]
OPTIONS = vsplat.TrainerOptions(force_fixed_point=True, sh_storage=vsplat.SH_BF16, tile_shape=vsplat.TILES_AUTO)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1)


def _view_pass(cloud, cam, target):
    m = O.Model(cloud, cam, sh_storage=vsplat.SH_BF16)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_fixed(r)
    return m, r


def _model_batch(cloud, cams, targets):
    """Oracle B per view (the passes in parallel threads: ctypes lets go of the
    GIL), then the chain rules in view order, gradients added in float32."""
    with ThreadPoolExecutor(max_workers=len(cams)) as pool:
        passes = list(pool.map(lambda ct: _view_pass(cloud, *ct), zip(cams, targets)))
    total, stats, counts = None, None, []
    for m, r in passes:
        st = np.zeros((cloud.count, 8), np.float32) if stats is None else stats
        g, stats = m.gradients(r, stats=st)
        g = [np.asarray(x, np.float32).reshape(-1) for x in g]
        total = [x.copy() for x in g] if total is None else [a + b for a, b in zip(total, g)]
        counts.append((r.visible, r.entries))
    return total, counts, passes[0][0].words


def _size_entries(t, cams, targets):
    """A synchronous batch in gradients mode grows the reservation to fit every
    view (and changes no parameter)."""
    t.step_views(cams, targets)
    t.reset_densify_stats()


@pytest.mark.parametrize("schedule", [vsplat.BATCH_VIEW_STREAMS, vsplat.BATCH_ONE_LAUNCH])
def test_config3_captured_step_bit_exact_at_20k(schedule):
    w = F.config(3, count=20_000)
    assert [(c.width, c.height) for c in w.cameras] == SIZES
    cloud, cams, targets = w.cloud, w.cameras, w.targets
    want_g, counts, words = _model_batch(cloud, cams, targets)
    pos, sc, rot, op, wd = O.adam_from_gradients(cloud, words, want_g, RATE_TUPLE, rest_scale=RATES.sh_rest_scale,
                                                 sh_storage=vsplat.SH_BF16)
    buffers = [vsplat.DeviceBuffer(x) for x in targets]
    try:
        with vsplat.Trainer(cloud, dataclasses.replace(OPTIONS, batch_raster=schedule)) as t:
            shapes = t.autotune_views(cams, [d.ptr for d in buffers])
            print(f"tile shapes: {shapes}", flush=True)
            _size_entries(t, cams, targets)
            t.capture_views(cams, [d.ptr for d in buffers], vsplat.MODE_STEP, RATES)
            t.launch_captured(1)
            stats = t.synchronize_views()
            c = t.cloud()
    finally:
        for d in buffers:
            d.free()
    # Oracle B lists 16x16 tiles; the entry count of a view the tuner gave
    # another shape differs by construction (its sums and steps do not).
    assert [s.visible_splats for s in stats] == [v for v, _ in counts]
    for shape, s, (_, entries) in zip(shapes, stats, counts):
        if tuple(shape) == (16, 16):
            assert s.tile_entries == entries
    np.testing.assert_array_equal(_bits(c.positions), _bits(pos))
    np.testing.assert_array_equal(_bits(c.scales), _bits(sc))
    np.testing.assert_array_equal(_bits(c.rotations), _bits(rot))
    np.testing.assert_array_equal(_bits(c.opacities), _bits(op))
    np.testing.assert_array_equal(_bits(c.sh), _bits(F.unpack_sh(wd, 16, vsplat.SH_BF16)))


def _check_lists(t, view, count):
    keys = t.inspect("entry_keys", view)
    values = t.inspect("entry_values", view)
    ranges = t.inspect("ranges", view)
    order = t.inspect("values", view)  # Gaussians in depth order
    depth_keys = t.inspect("keys", view)
    assert np.all(np.diff(depth_keys.astype(np.int64)) >= 0)
    rank = np.empty(count, np.int64)
    rank[order] = np.arange(count)
    assert np.all(np.diff(keys.astype(np.int64)) >= 0), "entries not sorted by tile"
    same = np.diff(keys.astype(np.int64)) == 0
    assert np.all(np.diff(rank[values])[same] > 0), "a tile's entries are not in depth order"
    used = ranges[ranges[:, 1] > ranges[:, 0]]
    assert used[:, 0].min() == 0 and used[:, 1].max() == keys.size
    assert int((ranges[:, 1] - ranges[:, 0]).sum()) == keys.size
    tiles = np.flatnonzero(ranges[:, 1] > ranges[:, 0])
    for tile in tiles[:: max(1, tiles.size // 64)]:
        a, b = ranges[tile]
        assert np.all(keys[a:b] == tile)


def test_config3_full_size_captured_properties():
    w = F.config(3)
    cloud, cams, targets = w.cloud, w.cameras, w.targets
    buffers = [vsplat.DeviceBuffer(x) for x in targets]
    ptrs = [d.ptr for d in buffers]
    clouds, stats_seen = [], []
    try:
        for run in range(2):
            with vsplat.Trainer(cloud, OPTIONS) as t:
                tuned = t.autotune_views(cams, ptrs)
                shapes = tuned  # the last run's, which `first` below belongs to
                _size_entries(t, cams, targets)
                t.capture_views(cams, ptrs, vsplat.MODE_STEP, RATES)
                t.launch_captured(1)
                first = t.synchronize_views()
                if run == 0:
                    print(f"tile shapes {tuned}; entries {[s.tile_entries for s in first]}", flush=True)
                    for k in range(len(cams)):
                        _check_lists(t, k, cloud.count)
                t.launch_captured(1)
                # Entry counts depend on the tile shape the tuner kept, which may
                # differ between the runs; images, sums and steps do not.
                stats_seen.append([(s.visible_splats, s.loss, s.l1, s.ssim) for s in t.synchronize_views()])
                clouds.append(t.cloud())
        a, b = clouds
        for x, y in zip((a.positions, a.scales, a.rotations, a.opacities, a.sh),
                        (b.positions, b.scales, b.rotations, b.opacities, b.sh)):
            np.testing.assert_array_equal(_bits(x), _bits(y))
        assert stats_seen[0] == stats_seen[1]
        assert np.any(_bits(a.positions) != _bits(cloud.positions))
        # Each view's counts in the batch (the last run's) equal a single-view
        # pass of the same view with the tile shape that run kept for it.
        for k in (0, 6):
            shape = {(16, 16): vsplat.TILES_16X16, (8, 8): vsplat.TILES_8X8, (32, 8): vsplat.TILES_32X8,
                     (32, 32): vsplat.TILES_32X32}[tuple(shapes[k])]
            with vsplat.Trainer(cloud, dataclasses.replace(OPTIONS, tile_shape=shape)) as single:
                s = single.compute_gradients(cams[k], targets[k])
            assert (s.visible_splats, s.tile_entries) == (first[k].visible_splats, first[k].tile_entries)
    finally:
        for d in buffers:
            d.free()
