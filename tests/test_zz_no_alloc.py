"""vsplat_render_noalloc, the render that allocates nothing (no counterpart in
the reference; for callers that cannot allocate GPU memory): the image it
writes into caller-owned GPU memory, with the library's static scratch, equals
vsplat_render's bit for bit, on calls in a row whose second is smaller, across
two trainers, and after a view that does not fit; sizes past the limits are
refused."""
import numpy as np
import pytest

from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

UP = (0.0, -1.0, 0.0)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1)


def _poisoned(pixels):
    return vsplat.DeviceBuffer(np.full(pixels * 4, np.nan, np.float32))


def _check(t, cam, out):
    print(f"render_noalloc {cam.width}x{cam.height}", flush=True)
    t.render_noalloc(cam, out)
    got = out.read(np.float32)[: cam.width * cam.height * 4]
    want = t.render(cam)
    np.testing.assert_array_equal(_bits(got), _bits(want))


def test_noalloc_render_equals_render_on_calls_in_a_row():
    cloud = F.config(1, count=20_000).cloud
    big = F.look_at((0, 0, -3), (0, 0, 0), UP, 60.0, 1280, 720)
    small = F.look_at((1.0, 0.5, -2.8), (0, 0, 0), UP, 60.0, 640, 480)
    out = _poisoned(big.width * big.height)
    try:
        with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
            _check(t, big, out)
            _check(t, small, out)  # the scratch still holds the larger view's lists
            _check(t, big, out)
        # Another trainer on the same GPU: another count, degree and tile shape.
        other = F.random_cloud(300, 1, 7)
        with vsplat.Trainer(other, vsplat.TrainerOptions(tile_shape=vsplat.TILES_8X8)) as t:
            _check(t, F.train_camera(100, 75), out)
    finally:
        out.free()


def test_noalloc_render_refuses_sizes_past_its_limits():
    out = vsplat.DeviceBuffer.empty(64 * 48 * 16)
    try:
        with vsplat.Trainer(F.random_cloud(60, 1, 13)) as t:
            with pytest.raises(ValueError, match="width\\*height\\*16"):
                t.render_noalloc(F.train_camera(72, 56), out)  # the output is too small
            with pytest.raises(ValueError, match="pixels"):
                t.render_noalloc(F.train_camera(2048, 1100), out)
            t.render_noalloc(F.train_camera(64, 48), out)
        with vsplat.Trainer(F.Cloud.zeros(vsplat.NOALLOC_MAX_GAUSSIANS + 1, 0)) as t:
            with pytest.raises(ValueError, match="Gaussians"):
                t.render_noalloc(F.train_camera(64, 48), out)
    finally:
        out.free()


def test_noalloc_render_reports_a_view_that_does_not_fit():
    w = F.config(4, count=100_000)  # heavy-tailed scales: far more than 2^21 entries at 1920x1080
    cam = w.cameras[0]
    small = F.look_at((0, 0, -3), (0, 0, 0), UP, 60.0, 160, 120)
    out = _poisoned(cam.width * cam.height)
    try:
        with vsplat.Trainer(w.cloud) as t:
            with pytest.raises(vsplat.BudgetExceeded, match="tile entries"):
                t.render_noalloc(cam, out)
            _check(t, small, out)  # the counters start clear again
    finally:
        out.free()
