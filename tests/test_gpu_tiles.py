"""Raster tile shapes (no reference counterpart; the reference's is 16x16):
8x8, 32x8, 32x32 and the autotuner's choice, on the GPU.

Each wave holds the same 8x8 block of pixels whatever the tile, and a tile's
list holds its splats in depth order, so the image, the fixed-point sums, the
gradients and the statistics are bit for bit those of 16x16 tiles (which are
pinned against Oracle B in test_gpu_parity.py). The lists themselves differ;
they are checked for what fixes them: sorted by tile, each tile's entries in
depth order, the ranges partitioning them, and every pixel a splat reaches
(brute force in double) lying in a tile listed for it.
"""
import math

import numpy as np
import pytest

from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

SHAPES = {vsplat.TILES_8X8: (8, 8), vsplat.TILES_32X8: (32, 8), vsplat.TILES_32X32: (32, 32)}


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _case(name):
    if name == "ref300":
        return F.random_cloud(300, 1, 7), F.train_camera(100, 75), F.smooth_target(100, 75, 23)
    if name == "deg3":
        return F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    if name == "config1_small":
        w = F.config(1, count=20_000)
        return w.cloud, w.cameras[0], w.targets[0]
    raise KeyError(name)


def _run(cloud, cam, target, tile_shape):
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, tile_shape=tile_shape)) as t:
        stats = t.compute_gradients(cam, target)
        out = {name: t.inspect(name) for name in ("image", "raw", "keys", "values", "entry_keys", "entry_blocks", "entry_values", "ranges",
                                                  "splats")}
        g = t.gradients()
        out["gradients"] = [g.positions, g.scales, g.rotations, g.opacities, g.sh]
        out["stats"] = t.densify_stats()
    return stats, out


@pytest.mark.parametrize("name", ["ref300", "deg3", "config1_small"])
@pytest.mark.parametrize("shape", [vsplat.TILES_8X8, vsplat.TILES_32X8, vsplat.TILES_32X32])
def test_tile_shape_gives_the_same_results(name, shape):
    cloud, cam, target = _case(name)
    base_stats, base = _run(cloud, cam, target, vsplat.TILES_16X16)
    stats, got = _run(cloud, cam, target, shape)
    print(f"{name} {SHAPES[shape]}: entries {stats.tile_entries} (16x16: {base_stats.tile_entries})", flush=True)
    assert stats.visible_splats == base_stats.visible_splats
    assert stats.loss == base_stats.loss and stats.l1 == base_stats.l1 and stats.ssim == base_stats.ssim
    np.testing.assert_array_equal(_bits(got["image"]), _bits(base["image"]))
    np.testing.assert_array_equal(got["raw"], base["raw"])
    for a, b in zip(got["gradients"], base["gradients"]):
        np.testing.assert_array_equal(_bits(a), _bits(b))
    np.testing.assert_array_equal(_bits(got["stats"]), _bits(base["stats"]))

    _check_lists(cloud, cam, got, stats, SHAPES[shape])
    _check_lists(cloud, cam, base, base_stats, (16, 16))



def _heavy_first(work):
    v = work.astype(np.int64) + 1
    e = np.floor(np.log2(v)).astype(np.int64)
    fraction = np.where(e >= 3, v >> np.maximum(e - 3, 0), v << np.maximum(3 - e, 0)) & 7
    return 255 - (8 * e + fraction)


def _check_order(order, keys):
    """A permutation of the launch indices that keeps each on its XCD (index
    mod 8) and runs each XCD's by non-decreasing key."""
    items = len(keys)
    assert len(order) == items
    assert sorted(order.tolist()) == list(range(items))
    assert (order % 8 == np.arange(items) % 8).all()
    assert (np.diff(keys[order].reshape(-1, 8), axis=0) >= 0).all()


@pytest.mark.parametrize("shape,large", [(vsplat.TILES_16X16, False), (vsplat.TILES_8X8, False), (vsplat.TILES_32X8, False),
                                         (vsplat.TILES_32X32, False), (vsplat.TILES_8X8, True)])
def test_raster_passes_run_each_xcds_blocks_heaviest_first(shape, large):
    """The raster passes' launch orders. The backward's runs each XCD's blocks
    by non-increasing work, the work of a block being the entries from its
    pixels' lowest stop to its tile list's end, recomputed here from the ranges
    and stops; the forward's by non-increasing length of the block's tile list.
    The large case (a 3072x3072 view of 8x8 tiles: 18432 blocks per XCD) has
    more blocks per XCD than the order kernels keep in LDS between their passes."""
    w = F.config(1, count=20_000)
    cam, target = w.cameras[0], w.targets[0]
    if large:
        cam = F.look_at((0, 0, -3), (0, 0, 0), (0.0, -1.0, 0.0), 60.0, 3072, 3072)
        target = F.smooth_target(3072, 3072, 29)
    tw, th = SHAPES.get(shape, (16, 16))
    with vsplat.Trainer(w.cloud, vsplat.TrainerOptions(force_fixed_point=True, tile_shape=shape)) as t:
        t.compute_gradients(cam, target)
        work, order, list_order = t.inspect("block_work"), t.inspect("block_order"), t.inspect("list_order")
        ranges, stops = t.inspect("ranges"), t.inspect("stops").reshape(cam.height, cam.width)
    grid_w, grid_h = (cam.width + tw - 1) // tw, (cam.height + th - 1) // th
    tiles, across, per_tile = grid_w * grid_h, tw // 8, (tw // 8) * (th // 8)
    items = (tiles + 7) // 8 * 8 * per_tile
    assert len(work) == items
    expected = np.zeros(items, np.int64)
    length = np.zeros(items, np.int64)
    for h in range(items):
        u = h // 8
        tile, q = (u // per_tile) * 8 + h % 8, u % per_tile
        if tile < tiles:
            length[h] = int(ranges[tile][1]) - int(ranges[tile][0])
        x0, y0 = (tile % grid_w) * tw + (q % across) * 8, (tile // grid_w) * th + (q // across) * 8
        if tile >= tiles or x0 >= cam.width or y0 >= cam.height:
            continue
        end = int(ranges[tile][1])
        expected[h] = end - min(int(stops[y0:y0 + 8, x0:x0 + 8].min()), end)
    np.testing.assert_array_equal(work, expected)
    _check_order(order, _heavy_first(work))
    _check_order(list_order, _heavy_first(length))
    print(f"{tw}x{th}: {items} blocks, heaviest {work.max()} entries, median {int(np.median(work))}", flush=True)

def _check_lists(cloud, cam, got, stats, shape):
    # The lists: sorted by tile, depth order within a tile, ranges partitioning them.
    tw, th = shape
    grid_w, grid_h = (cam.width + tw - 1) // tw, (cam.height + th - 1) // th
    keys, values = got["entry_keys"][: stats.tile_entries], got["entry_values"][: stats.tile_entries]
    blocks = got["entry_blocks"][: stats.tile_entries]
    assert (np.diff(keys.astype(np.int64)) >= 0).all() and keys.max() < grid_w * grid_h
    assert blocks.max() < (1 << ((tw // 8) * (th // 8)))
    rank = np.empty(cloud.count, np.int64)
    rank[got["values"]] = np.arange(cloud.count)
    same_tile = np.diff(keys.astype(np.int64)) == 0
    assert (np.diff(rank[values])[same_tile] > 0).all()
    ranges = got["ranges"][: grid_w * grid_h]
    total = 0
    for t in range(grid_w * grid_h):
        a, b = ranges[t]
        if b > a:
            assert (keys[a:b] == t).all()
            total += b - a
    assert total == stats.tile_entries

    # Brute force: every pixel where a splat's alpha reaches 1/255 lies in a
    # tile listed for it, and in a block of that tile its entry marks.
    listed = {}
    for key, value, mask in zip(keys, values, blocks):
        listed.setdefault(int(value), {})[int(key)] = int(mask)
    splats = got["splats"]
    visible = got["values"][got["keys"] != 0xFFFFFFFF]
    rng = np.random.default_rng(0)
    for i in rng.choice(visible, min(len(visible), 60), replace=False):
        s = splats[i].astype(np.float64)
        radius = int(s[7]) + 1
        cx, cy = int(s[8]), int(s[9])
        for y in range(max(0, cy - radius), min(cam.height, cy + radius + 1)):
            for x in range(max(0, cx - radius), min(cam.width, cx + radius + 1)):
                dx, dy = x + 0.5 - s[8], y + 0.5 - s[9]
                if abs(dx) > s[7] or abs(dy) > s[7]:
                    continue
                power = -0.5 * (s[0] * dx * dx + s[2] * dy * dy) - s[1] * dx * dy
                alpha = min(0.99, s[3] * math.exp(power)) if power <= 0 else 0.0
                if alpha >= (1 / 255) * (1 + 1e-3):
                    tile = (y // th) * grid_w + x // tw
                    assert tile in listed.get(int(i), {}), (i, x, y)
                    block = ((y % th) // 8) * (tw // 8) + (x % tw) // 8
                    assert (listed[int(i)][tile] >> block) & 1, (i, x, y, tile, block)


def test_autotuned_shape_steps_as_16x16():
    cloud, cam, target = _case("config1_small")
    rates = vsplat.LearningRates()
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as a, \
         vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, tile_shape=vsplat.TILES_AUTO)) as b, \
         vsplat.DeviceBuffer(target) as d:
        shape = b.autotune_tiles(cam, d.ptr, 2)
        print(f"autotuned {shape} for {cam.width}x{cam.height}", flush=True)
        assert shape in ((16, 16), (8, 8), (32, 8), (32, 32))
        for _ in range(2):
            a.step(cam, target, rates)
            b.step(cam, target, rates)
        ca, cb = a.cloud(), b.cloud()
    for x, y in zip((ca.positions, ca.scales, ca.rotations, ca.opacities, ca.sh),
                    (cb.positions, cb.scales, cb.rotations, cb.opacities, cb.sh)):
        np.testing.assert_array_equal(_bits(x), _bits(y))


def test_tuner_leaves_the_entry_buffers_fit_for_its_choice():
    """The tuner tries 8x8 tiles, which list several times the entries of the
    others; after it, the tile-entry buffers are at most twice what the chosen
    shape's entries take by the growth rule (the sort's grids are sized by
    them), and the trainer's memory is that of a trainer given the shape
    (whichever it chose)."""
    cloud, cam, target = _case("config1_small")
    named = {(16, 16): vsplat.TILES_16X16, (8, 8): vsplat.TILES_8X8, (32, 8): vsplat.TILES_32X8, (32, 32): vsplat.TILES_32X32}
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, tile_shape=vsplat.TILES_AUTO)) as b, \
         vsplat.DeviceBuffer(target) as d:
        shape = b.autotune_tiles(cam, d.ptr, 1)
        b.compute_gradients(cam, target)
        tuned = b.gpu_bytes()
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, tile_shape=named[tuple(shape)])) as a:
        stats = a.compute_gradients(cam, target)
        given = a.gpu_bytes()
        with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, tile_shape=vsplat.TILES_8X8)) as e:
            eight = e.compute_gradients(cam, target).tile_entries
    wanted = stats.tile_entries + stats.tile_entries // 2 + 1024
    print(f"chose {shape}: {stats.tile_entries} entries (8x8: {eight}); tuned {tuned} B, given {given} B", flush=True)
    if tuple(shape) != (8, 8):  # the tuner may pick 8x8 itself on this small cloud
        assert eight > 2 * stats.tile_entries
    assert tuned <= given + 2 * wanted * vsplat.BYTES_PER_TILE_ENTRY


def test_batch_autotuner_keeps_the_batch_results():
    """The batch tuner (vsplat_autotune_views) picks a shape per view size by
    timing the whole batch, changes no parameter, gradient or statistic, and
    the batch's gradient sums then equal those of 16x16 tiles bit for bit."""
    w = F.config(3, count=20_000)
    cams, targets = w.cameras[:4], w.targets[:4]
    buffers = [vsplat.DeviceBuffer(t) for t in targets]
    try:
        with vsplat.Trainer(w.cloud, vsplat.TrainerOptions(force_fixed_point=True)) as a, \
             vsplat.Trainer(w.cloud, vsplat.TrainerOptions(force_fixed_point=True, tile_shape=vsplat.TILES_AUTO)) as b:
            shapes = b.autotune_views(cams, [x.ptr for x in buffers], 2)
            print(f"batch shapes {shapes}", flush=True)
            assert len(shapes) == len(cams) and all(s in ((16, 16), (8, 8), (32, 8), (32, 32)) for s in shapes)
            for cam, shape in zip(cams, shapes):  # one shape per size
                assert all(t == shape for c, t in zip(cams, shapes) if (c.width, c.height) == (cam.width, cam.height))
            assert not np.any(b.densify_stats())
            a.step_views(cams, targets)
            b.step_views(cams, targets)
            ga, gb = a.gradients(), b.gradients()
            sa, sb = a.densify_stats(), b.densify_stats()
    finally:
        for x in buffers:
            x.free()
    for name in ("positions", "scales", "rotations", "opacities", "sh"):
        np.testing.assert_array_equal(_bits(getattr(ga, name)), _bits(getattr(gb, name)))
    np.testing.assert_array_equal(_bits(sa), _bits(sb))
