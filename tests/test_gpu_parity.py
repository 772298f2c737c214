"""GPU parity: the HIP path through the C interface against Oracle B (the FP32
model of the GPU algorithm), stage by stage and bit for bit, and against Oracle A
(the FP64 restatement of the reference's src/train/reference.cpp) within the
reference's own tolerances (tests/test_train.cpp:218-293, 353-403).

Bit-exact: device exp/log; projected splats; depth-sorted keys and values; tile
coverage, entries (after the stable tile sort) and ranges; the image, log
transmittance and stop of every pixel; the SSIM fields and the adjoint; the
14 fixed-point sums per Gaussian; and the gradients the chain rule derives from them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "ref60":  # test_train.cpp:256-293
        return F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.smooth_target(72, 56, 11)
    if name == "ref300":  # test_train.cpp:218-254
        return F.random_cloud(300, 1, 7), F.train_camera(100, 75), F.smooth_target(100, 75, 5)
    if name == "stop400":  # test_train.cpp:353-403: nearly opaque, most pixels stop early
        c = F.random_cloud(400, 1, 17)
        c.opacities[:] = 5.0
        return c, F.train_camera(80, 60), F.smooth_target(80, 60, 19)
    if name == "deg3":
        return F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    if name == "config0":
        w = F.config(0)
        return w.cloud, w.cameras[0], w.targets[0]
    if name == "config1_small":  # the 1280x720 workload's distributions at 20k Gaussians
        w = F.config(1, count=20_000)
        return w.cloud, w.cameras[0], w.targets[0]
    if name == "unseen":  # the camera looks away from every Gaussian: empty lists, the background, untouched Gaussians
        return F.random_cloud(60, 1, 13), F.look_at((0, 0, -3), (0, 0, -6), (0.0, -1.0, 0.0), 60.0, 72, 56), F.smooth_target(72, 56, 11)
    if name == "heavy_tails":  # the stress workload's distributions at 3000 Gaussians, 640x480: many boxes of more than 64 tiles
        w = F.config(4, count=3000)
        cam = F.look_at((0, 0, -3), (0, 0, 0), (0.0, -1.0, 0.0), 60.0, 640, 480)
        return w.cloud, cam, F.smooth_target_explicit(640, 480, 1004)
    raise KeyError(name)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_device_math_is_bit_exact():
    rng = np.random.default_rng(1)
    words = rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, 1.17549435e-38, 88.7, 89.0, -87.3, -103.9, -104.1,
                        1.0, 0.99, 0.5, 1.0 / 255.0], np.float32)
    x = np.concatenate([words.view(np.float32), special])
    small = np.concatenate([rng.uniform(-20, 5, 1 << 18).astype(np.float32), rng.uniform(0, 1, 1 << 18).astype(np.float32)])
    for xs in (x, small):
        for fn, ref in ((0, O.exp32), (1, O.log32)):
            got, want = vsplat.math_probe(fn, xs), ref(xs)
            same = (_bits(got) == _bits(want)) | (np.isnan(got) & np.isnan(want))
            assert same.all(), f"function {fn}: {np.count_nonzero(~same)} differ, first at {xs[~same][:4]}"
    pos = np.abs(x[np.isfinite(x)])
    assert (_bits(vsplat.math_probe(2, pos)) == _bits(np.sqrt(pos))).all()
    nz = x[np.isfinite(x) & (x != 0)]
    with np.errstate(over="ignore", divide="ignore"):
        assert (_bits(vsplat.math_probe(3, nz)) == _bits(np.float32(1.0) / nz)).all()
    fin = x[np.isfinite(x)]
    with np.errstate(over="ignore"):
        assert (_bits(vsplat.math_probe(4, fin)) == _bits(fin.astype(np.float16).astype(np.float32))).all()
    halves = np.arange(65536, dtype=np.uint32)
    got = vsplat.math_probe(5, halves.view(np.float32))
    want = halves.astype(np.uint16).view(np.float16).astype(np.float32)
    assert ((_bits(got) == _bits(want)) | (np.isnan(got) & np.isnan(want))).all()


def test_backward_division_is_ieee_division():
    """The backward's division without the hardware's range scaling equals float32 a / b over the
    range it is used in (IEEE division rounds correctly, numpy's too)."""
    rng = np.random.default_rng(2)
    m = 1 << 20
    # The backward's own operands (transmittance in [1e-4, 1], 1 - alpha in [0.01, 1]), then the
    # whole stated range, then the ends and mantissas of all ones next to ones.
    a = [rng.uniform(1e-4, 1.0, m), 2.0 ** rng.uniform(-100, 60, m)]
    b = [rng.uniform(0.01, 1.0, m), 2.0 ** rng.uniform(-20, 0, m)]
    ends_a = np.array([2.0 ** -100, 2.0 ** 60, 1e-4, 0.5, 1.0 - 2.0 ** -24, 1.0, 1.0 + 2.0 ** -23, 3.0, 0.1], np.float32)
    ends_b = np.array([2.0 ** -20, 0.01, np.nextafter(np.float32(0.01), np.float32(1)), 1.0 - 1.0 / 255.0, 0.75,
                       1.0 - 2.0 ** -23, 1.0 - 2.0 ** -24, 1.0], np.float32)
    ea, eb = np.meshgrid(ends_a, ends_b)
    a = np.concatenate(a + [ea.ravel()]).astype(np.float32)
    b = np.concatenate(b + [eb.ravel()]).astype(np.float32)
    got = vsplat.math_probe(6, np.stack([a, b], 1).ravel())
    want = a / b
    same = _bits(got[0::2]) == _bits(want)
    assert same.all(), f"{np.count_nonzero(~same)} differ, first at a={a[~same][:4]} b={b[~same][:4]}"
    assert (got[1::2] == 0).all()


def _run_both(name, fixed=True, lam=0.2, sh_storage=0):
    cloud, cam, target = _case(name)
    model = O.Model(cloud, cam, sh_storage=sh_storage)
    r = model.forward()
    model.loss(r, target, lam)
    if fixed:
        model.backward_fixed(r)
    else:
        model.backward_float(r)
    g_model, st_model = model.gradients(r)
    t = vsplat.Trainer(cloud, vsplat.TrainerOptions(lambda_ssim=lam, force_fixed_point=fixed, sh_storage=sh_storage))
    stats = t.compute_gradients(cam, target)
    return cloud, cam, target, model, r, g_model, st_model, t, stats


@pytest.mark.parametrize("name", ["ref60", "ref300", "stop400", "deg3", "config0", "config1_small", "heavy_tails", "unseen"])
def test_pipeline_bit_exact_fixed_point(name):
    cloud, cam, target, model, r, g_model, st_model, t, stats = _run_both(name, fixed=True)
    print(f"{name}: visible {r.visible} entries {r.entries}")
    if name == "unseen":
        assert r.visible == 0 and r.entries == 0
    if name == "heavy_tails":  # splats listed by the big-splat kernels are in the case
        s = r.splats[r.values[r.keys != 0xFFFFFFFF]]
        tiles_x = np.floor((s[:, 8] + s[:, 10] + 1) / 16) - np.floor((s[:, 8] - s[:, 10] - 1) / 16) + 1
        tiles_y = np.floor((s[:, 9] + s[:, 11] + 1) / 16) - np.floor((s[:, 9] - s[:, 11] - 1) / 16) + 1
        assert np.count_nonzero(tiles_x * tiles_y > 64) >= 20
    assert stats.visible_splats == r.visible
    assert stats.tile_entries == r.entries
    np.testing.assert_array_equal(_bits(t.inspect("keys")), _bits(r.keys))
    np.testing.assert_array_equal(t.inspect("values"), r.values)
    visible = r.values[r.keys != 0xFFFFFFFF]
    np.testing.assert_array_equal(_bits(t.inspect("splats")[visible]), _bits(r.splats[visible]))
    np.testing.assert_array_equal(t.inspect("entry_keys"), r.entry_keys)
    np.testing.assert_array_equal(t.inspect("entry_values"), r.entry_values)
    np.testing.assert_array_equal(t.inspect("ranges"), r.ranges)
    np.testing.assert_array_equal(_bits(t.inspect("image")), _bits(r.image))
    np.testing.assert_array_equal(t.inspect("stops"), r.stops)
    np.testing.assert_array_equal(_bits(t.inspect("fields")), _bits(r.fields))
    np.testing.assert_array_equal(_bits(t.inspect("adjoint")), _bits(r.adjoint))
    np.testing.assert_array_equal(_bits(t.inspect("tile_totals")), _bits(r.tile_totals))
    np.testing.assert_array_equal(t.inspect("raw"), r.raw_fixed)
    g = t.gradients()
    for got, want in zip((g.positions, g.scales, g.rotations, g.opacities, g.sh), g_model):
        np.testing.assert_array_equal(_bits(got.reshape(-1)), _bits(want))
    np.testing.assert_array_equal(_bits(t.densify_stats()), _bits(st_model))


@pytest.mark.parametrize("name", ["deg3", "config1_small"])
def test_pipeline_bit_exact_with_bf16_sh(name):
    """SH stored as bf16: splats, image, raw sums and gradients bit for bit
    against Oracle B with the same storage."""
    cloud, cam, target, model, r, g_model, st_model, t, stats = _run_both(name, fixed=True, sh_storage=vsplat.SH_BF16)
    assert stats.visible_splats == r.visible and stats.tile_entries == r.entries
    visible = r.values[r.keys != 0xFFFFFFFF]
    np.testing.assert_array_equal(_bits(t.inspect("splats")[visible]), _bits(r.splats[visible]))
    np.testing.assert_array_equal(_bits(t.inspect("image")), _bits(r.image))
    np.testing.assert_array_equal(t.inspect("raw"), r.raw_fixed)
    g = t.gradients()
    for got, want in zip((g.positions, g.scales, g.rotations, g.opacities, g.sh), g_model):
        np.testing.assert_array_equal(_bits(got.reshape(-1)), _bits(want))
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as half:
        assert t.bytes_per_gaussian() == half.bytes_per_gaussian() - 8  # 24 SH words instead of 26


def _check_close(a, b, what, min_cos, max_ratio):
    """check_close of tests/test_train.cpp:38-53: direction and length."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    assert (a @ a) > 0 and (b @ b) > 0, what
    cos = (a @ b) / np.sqrt((a @ a) * (b @ b))
    ratio = np.sqrt((a @ a) / (b @ b))
    print(f"  {what:10s} cosine {cos:.6f} ratio {ratio:.5f}")
    assert cos > min_cos, what
    assert abs(ratio - 1) < max_ratio, what


@pytest.mark.parametrize("name", ["ref60", "stop400", "deg3"])
@pytest.mark.parametrize("fixed", [False, True])
def test_against_fp64_reference(name, fixed):
    """tests/test_train.cpp:256-293 and 353-403: forward <= 1e-3, log T <= 1e-2,
    loss/L1/SSIM <= 1e-3; gradients cosine > 0.999 and length within 2% with
    float sums, > 0.97 and 10% with fixed point (the reference's bars)."""
    cloud, cam, target = _case(name)
    ref = O.render(cloud, cam)
    ref_loss = O.loss(ref, target, 0.2)
    ref_g = O.gradients(cloud, cam, ref_loss.adjoint)
    t = vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=fixed))
    image = t.render(cam)
    assert np.abs(image[:, :3] - ref.rgb).max() < 1e-3
    assert np.abs(image[:, 3] - ref.log_t).max() < 1e-2
    stats = t.compute_gradients(cam, target)
    assert abs(stats.loss - ref_loss.loss) < 1e-3
    assert abs(stats.l1 - ref_loss.l1) < 1e-3
    assert abs(stats.ssim - ref_loss.ssim) < 1e-3
    g = t.gradients()
    cos, ratio = (0.97, 0.1) if fixed else (0.999, 0.02)
    for got, want, what in ((g.positions, ref_g.positions, "positions"), (g.scales, ref_g.scales, "scales"),
                            (g.rotations, ref_g.rotations, "rotations"), (g.opacities, ref_g.opacities, "opacities"),
                            (g.sh, ref_g.sh, "sh")):
        _check_close(got, want, what, cos, ratio)


def test_needle_splats_bit_exact():
    """Long, thin splats whose 2D covariance is built and inverted in double:
    splats, image and fixed-point sums bit for bit against Oracle B."""
    from test_oracle_b import _needle_scene
    cloud, cam = _needle_scene(60.0, 0.03), F.train_camera(72, 56)
    target = F.smooth_target(72, 56, 5)
    model = O.Model(cloud, cam)
    r = model.forward()
    model.loss(r, target, 0.2)
    model.backward_fixed(r)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        t.compute_gradients(cam, target)
        visible = r.values[r.keys != 0xFFFFFFFF]
        assert {30, 31, 32} <= set(visible.tolist())
        np.testing.assert_array_equal(_bits(t.inspect("splats")[visible]), _bits(r.splats[visible]))
        np.testing.assert_array_equal(_bits(t.inspect("image")), _bits(r.image))
        np.testing.assert_array_equal(t.inspect("raw"), r.raw_fixed)
