"""The batched step (several views per iteration; no reference counterpart) on
the GPU against Oracle B:

* gradients summed over views of different sizes, bit for bit: each view's
  gradients from the FP32 model, added in view order in float32; the
  densification statistics likewise; each view's visible count, entries and
  loss;
* the same over more views than one chain-rule launch takes;
* a batch of one view steps exactly as vsplat_step does;
* a batch of two views steps as Oracle B's Adam from the summed gradients;
* the queued form with device targets gives the same result;
* a batch captured as a HIP graph and replayed steps as the queued form does,
  replay after replay, and refuses to replay once a buffer it uses is remade.

The sums, steps and captured replays are checked under both batch schedules:
each view's passes on a stream of its own, and one forward launch and one
backward launch over every view's 8x8 blocks (batch_raster).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

RATES = vsplat.LearningRates(position=1.6e-4, scale=0.005, rotation=0.001, opacity=0.05, sh=0.025, sh_rest_scale=0.05)
RATE_TUPLE = (RATES.position, RATES.scale, RATES.rotation, RATES.opacity, RATES.sh)
SIZES = [(64, 48), (96, 54), (80, 60)]


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _ring(count=2000, degree=3, seed=77, sizes=SIZES):
    """A small version of the eight-view ring: views of different sizes."""
    cloud = F.synthetic_cloud(count, degree, seed, sh_sigma=0.3)
    up = (0.0, -1.0, 0.0)
    cams, targets = [], []
    for k, (w, h) in enumerate(sizes):
        th = 2.0 * math.pi * k / len(sizes)
        cams.append(F.look_at((3.0 * math.sin(th), 0.0, -3.0 * math.cos(th)), (0, 0, 0), up, 60.0, w, h))
        targets.append(F.smooth_target_explicit(w, h, seed * 16 + k))
    return cloud, cams, targets


def _model_sum(cloud, cams, targets, sh_storage=0):
    """Oracle B per view (fixed point), gradients and statistics summed in view order."""
    total, stats, per_view = None, None, []
    for cam, target in zip(cams, targets):
        m = O.Model(cloud, cam, sh_storage=sh_storage)
        r = m.forward()
        m.loss(r, target, 0.2)
        m.backward_fixed(r)
        st = np.zeros((cloud.count, 8), np.float32) if stats is None else stats
        g, stats = m.gradients(r, stats=st)
        g = [np.asarray(x, np.float32).reshape(-1) for x in g]
        total = [x.copy() for x in g] if total is None else [a + b for a, b in zip(total, g)]
        tiles = r.tile_totals.astype(np.float64).sum(0) / (3.0 * cam.width * cam.height)
        per_view.append((r.visible, r.entries, float(tiles[1]), float(tiles[0])))
    return total, stats, per_view, m.words


SCHEDULES = [vsplat.BATCH_VIEW_STREAMS, vsplat.BATCH_ONE_LAUNCH]


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("tile_shape", [vsplat.TILES_16X16, vsplat.TILES_32X8, vsplat.TILES_32X32])
def test_gradients_summed_over_views_bit_exact(schedule, tile_shape):
    # Every tile shape gives the 16x16 sums (the batch's raster kernels read
    # the lists' block marks with 32x32 tiles); Oracle B lists 16x16 tiles, so
    # only those entry counts are compared.
    cloud, cams, targets = _ring()
    want, want_stats, per_view, _ = _model_sum(cloud, cams, targets)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, batch_raster=schedule, tile_shape=tile_shape)) as t:
        stats = t.step_views(cams, targets)
        g = t.gradients()
        dens = t.densify_stats()
    for got, w in zip((g.positions, g.scales, g.rotations, g.opacities, g.sh), want):
        np.testing.assert_array_equal(_bits(got.reshape(-1)), _bits(w))
    np.testing.assert_array_equal(_bits(dens), _bits(want_stats))
    assert len(stats) == len(cams)
    for s, (visible, entries, l1, ssim) in zip(stats, per_view):
        assert s.visible_splats == visible
        assert tile_shape != vsplat.TILES_16X16 or s.tile_entries == entries
        assert abs(s.l1 - l1) < 1e-6 and abs(s.ssim - ssim) < 1e-6


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_more_views_than_one_chain_launch_takes(schedule):
    """Nine views: the chain rule runs in two chunks (eight views, then one
    continuing the sums), and the one-launch schedule's raster launches take
    eight views each, which must not change the sums."""
    cloud, cams, targets = _ring(1500, 2, 91, sizes=[(48, 32), (64, 40), (40, 40)] * 3)
    want, want_stats, _, _ = _model_sum(cloud, cams, targets)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, batch_raster=schedule)) as t:
        t.step_views(cams, targets)
        g = t.gradients()
        dens = t.densify_stats()
    for got, w in zip((g.positions, g.scales, g.rotations, g.opacities, g.sh), want):
        np.testing.assert_array_equal(_bits(got.reshape(-1)), _bits(w))
    np.testing.assert_array_equal(_bits(dens), _bits(want_stats))


def test_batch_of_one_steps_as_step_does():
    cloud, cams, targets = _ring()
    options = vsplat.TrainerOptions(force_fixed_point=True)
    with vsplat.Trainer(cloud, options) as a, vsplat.Trainer(cloud, options) as b:
        a.step(cams[1], targets[1], RATES)
        b.step_views([cams[1]], [targets[1]], RATES)
        ca, cb = a.cloud(), b.cloud()
    for x, y in zip((ca.positions, ca.scales, ca.rotations, ca.opacities, ca.sh),
                    (cb.positions, cb.scales, cb.rotations, cb.opacities, cb.sh)):
        np.testing.assert_array_equal(_bits(x), _bits(y))


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("sh_storage", [0, 1])
def test_batch_of_two_steps_as_the_model(sh_storage, schedule):
    cloud, cams, targets = _ring()
    want_g, _, _, words = _model_sum(cloud, cams[:2], targets[:2], sh_storage)
    pos, sc, rot, op, w = O.adam_from_gradients(cloud, words, want_g, RATE_TUPLE, rest_scale=RATES.sh_rest_scale,
                                                sh_storage=sh_storage)
    options = vsplat.TrainerOptions(force_fixed_point=True, sh_storage=sh_storage, batch_raster=schedule)
    with vsplat.Trainer(cloud, options) as t:
        t.step_views(cams[:2], targets[:2], RATES)
        c = t.cloud()
    np.testing.assert_array_equal(_bits(c.positions.reshape(-1)), _bits(pos))
    np.testing.assert_array_equal(_bits(c.scales.reshape(-1)), _bits(sc))
    np.testing.assert_array_equal(_bits(c.rotations.reshape(-1)), _bits(rot))
    np.testing.assert_array_equal(_bits(c.opacities), _bits(op))
    sh = F.unpack_sh(w, (cloud.sh_degree + 1) ** 2, sh_storage).reshape(c.sh.shape)
    np.testing.assert_array_equal(_bits(c.sh), _bits(sh))


def test_queued_batch_matches_the_synchronous_one():
    cloud, cams, targets = _ring()
    options = vsplat.TrainerOptions(force_fixed_point=True)
    buffers = [vsplat.DeviceBuffer(x) for x in targets]
    try:
        with vsplat.Trainer(cloud, options) as a, vsplat.Trainer(cloud, options) as b:
            a.step_views(cams, targets, RATES)
            b.reserve_entries(1 << 16)
            b.enqueue_views(cams, [d.ptr for d in buffers], vsplat.MODE_STEP, RATES)
            stats = b.synchronize_views()
            assert len(stats) == len(cams)
            ca, cb = a.cloud(), b.cloud()
    finally:
        for d in buffers:
            d.free()
    for x, y in zip((ca.positions, ca.opacities, ca.sh), (cb.positions, cb.opacities, cb.sh)):
        np.testing.assert_array_equal(_bits(x), _bits(y))


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_captured_batch_replays_as_queued_steps(schedule):
    cloud, cams, targets = _ring()
    options = vsplat.TrainerOptions(force_fixed_point=True, batch_raster=schedule)
    buffers = [vsplat.DeviceBuffer(x) for x in targets]
    ptrs = [d.ptr for d in buffers]
    try:
        with vsplat.Trainer(cloud, options) as a, vsplat.Trainer(cloud, options) as b:
            for t in (a, b):
                t.reserve_entries(1 << 16)
            b.capture_views(cams, ptrs, vsplat.MODE_STEP, RATES)
            for replay in range(3):
                a.enqueue_views(cams, ptrs, vsplat.MODE_STEP, RATES)
                want = a.synchronize_views()
                b.launch_captured()
                got = b.synchronize_views()
                assert [s.tile_entries for s in got] == [s.tile_entries for s in want]
                assert [s.loss for s in got] == [s.loss for s in want]
                ca, cb = a.cloud(), b.cloud()
                for x, y in zip((ca.positions, ca.scales, ca.rotations, ca.opacities, ca.sh),
                                (cb.positions, cb.scales, cb.rotations, cb.opacities, cb.sh)):
                    np.testing.assert_array_equal(_bits(x), _bits(y), err_msg=f"replay {replay}")
            # Two replays queued back to back are two more steps.
            for _ in range(2):
                a.enqueue_views(cams, ptrs, vsplat.MODE_STEP, RATES)
            a.synchronize_views()
            b.launch_captured(2)
            b.synchronize_views()
            ca, cb = a.cloud(), b.cloud()
            np.testing.assert_array_equal(_bits(ca.positions), _bits(cb.positions))
            np.testing.assert_array_equal(_bits(ca.sh), _bits(cb.sh))
            # A larger view remakes the first view's per-pixel buffers: the graph is stale.
            big = F.look_at((0.0, 0.0, -3.0), (0, 0, 0), (0.0, -1.0, 0.0), 60.0, 160, 120)
            b.render(big)
            with pytest.raises(vsplat.LogicError, match="capture it again"):
                b.launch_captured()
    finally:
        for d in buffers:
            d.free()


def test_captured_batch_refuses_a_later_exposure():
    """An exposure set after the capture would be ignored by a replay (the
    batch takes the neutral one): the replay refuses instead."""
    cloud, cams, targets = _ring()
    buffers = [vsplat.DeviceBuffer(x) for x in targets]
    try:
        with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
            t.reserve_entries(1 << 16)
            t.capture_views(cams, [d.ptr for d in buffers], vsplat.MODE_STEP, RATES)
            t.set_exposure([1.1, 0, 0, 0.01, 0, 1, 0, 0, 0, 0, 1, 0])
            with pytest.raises(vsplat.LogicError):
                t.launch_captured()
            t.set_exposure(O.NEUTRAL_EXPOSURE)
            t.launch_captured()
            t.synchronize_views()
    finally:
        for d in buffers:
            d.free()
