"""Oracle B, the FP32 model of the GPU pipeline, on the CPU:

* against Oracle A (the FP64 reference restatement) with the reference's own
  tolerances (tests/test_train.cpp:218-293, 353-403): the model's arithmetic is
  this project's choice, the algorithm must still be the reference's;
* its integer stages against what fixes them independently: the tile lists
  cover every pixel a splat reaches (brute force over pixels, in double), the
  sort is a stable sort, the ranges partition the sorted entries;
* the reference's invariants: degenerate Gaussians do not change the picture
  (test_train.cpp:544-586); fixed-point sums equal the float sums to within
  their rounding; the first Adam step moves every parameter by its rate
  (test_train.cpp:450-491).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F


def check_close(a, b, what, min_cos, max_ratio):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    cos = (a @ b) / math.sqrt((a @ a) * (b @ b))
    ratio = math.sqrt((a @ a) / (b @ b))
    assert cos > min_cos, f"{what}: cosine {cos}"
    assert abs(ratio - 1) < max_ratio, f"{what}: ratio {ratio}"


def _case(name):
    if name == "ref60":
        return F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.smooth_target(72, 56, 11)
    if name == "ref300":
        return F.random_cloud(300, 1, 7), F.train_camera(100, 75), F.smooth_target(100, 75, 5)
    if name == "stop400":
        c = F.random_cloud(400, 1, 17)
        c.opacities[:] = 5.0
        return c, F.train_camera(80, 60), F.smooth_target(80, 60, 19)
    if name == "deg3":
        return F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    if name == "config0":
        w = F.config(0, count=3000)
        return w.cloud, w.cameras[0], w.targets[0]
    raise KeyError(name)


@pytest.mark.parametrize("name", ["ref60", "ref300", "stop400", "deg3"])
@pytest.mark.parametrize("fixed", [False, True])
def test_model_matches_fp64_reference(name, fixed):
    cloud, cam, target = _case(name)
    ref = O.render(cloud, cam, threads=2)
    m = O.Model(cloud, cam)
    r = m.forward()
    assert np.abs(r.image[:, :3] - ref.rgb).max() < 1e-3
    assert np.abs(r.image[:, 3] - ref.log_t).max() < 1e-2
    ref_loss = O.loss(ref, target, 0.2, threads=2)
    m.loss(r, target, 0.2)
    t = r.tile_totals.astype(np.float64).sum(0) / (3 * cam.width * cam.height)
    assert abs(t[1] - ref_loss.l1) < 1e-3 and abs(t[0] - ref_loss.ssim) < 1e-3
    ref_g = O.gradients(cloud, cam, ref_loss.adjoint, threads=2)
    (m.backward_fixed if fixed else m.backward_float)(r)
    g, stats = m.gradients(r)
    cos, ratio = (0.97, 0.1) if fixed else (0.999, 0.02)
    for got, want, what in zip(g, (ref_g.positions, ref_g.scales, ref_g.rotations, ref_g.opacities, ref_g.sh),
                               ("positions", "scales", "rotations", "opacities", "sh")):
        check_close(got, want, what, cos, ratio)
    # The densification statistics: contribution (sum T alpha in pixels) and |dL/dcentre|
    # (test_train.cpp:386-392).
    check_close(stats[:, 3], O.contributions(cloud, cam, threads=2), "contribution", cos, ratio)
    absolute = np.hypot(ref_g.absolute_centre[:, 0] * 0.5 * cam.width, ref_g.absolute_centre[:, 1] * 0.5 * cam.height)
    check_close(stats[:, 4], absolute, "absolute", cos, ratio)


def _alpha64(s, x, y):
    """splat_at in double from the model's splat record (conic, opacity, radius, centre)."""
    dx, dy = x + 0.5 - float(s[8]), y + 0.5 - float(s[9])
    if abs(dx) > s[7] or abs(dy) > s[7]:
        return 0.0
    power = -0.5 * (float(s[0]) * dx * dx + float(s[2]) * dy * dy) - float(s[1]) * dx * dy
    return min(0.99, float(s[3]) * math.exp(power)) if power <= 0 else 0.0


@pytest.mark.parametrize("name", ["ref300", "deg3", "config0"])
def test_tile_lists_cover_every_reached_pixel(name):
    """Brute force: every pixel where a splat's alpha reaches 1/255 (with a
    margin for the model's rounding) lies in a tile listed for that splat, and
    every listed tile is within the splat's bounding box of tiles."""
    cloud, cam, _ = _case(name)
    m = O.Model(cloud, cam)
    r = m.forward()
    listed = {}
    for key, value in zip(r.entry_keys, r.entry_values):
        listed.setdefault(int(value), set()).add(int(key))
    visible = r.values[r.keys != 0xFFFFFFFF]
    assert len(visible) == r.visible
    rng = np.random.default_rng(0)
    for i in rng.choice(visible, min(len(visible), 150), replace=False):
        s = r.splats[i]
        radius = int(s[7]) + 1
        cx, cy = int(s[8]), int(s[9])
        for y in range(max(0, cy - radius), min(cam.height, cy + radius + 1)):
            for x in range(max(0, cx - radius), min(cam.width, cx + radius + 1)):
                if _alpha64(s, x, y) >= (1 / 255) * (1 + 1e-3):
                    assert (y // 16) * m.grid_w + x // 16 in listed.get(int(i), set()), (i, x, y)
    assert sum(len(v) for v in listed.values()) == r.entries


@pytest.mark.parametrize("name", ["ref300", "config0"])
def test_sort_and_ranges(name):
    cloud, cam, _ = _case(name)
    m = O.Model(cloud, cam)
    r = m.forward()
    # Depth sort: keys ascending (far first), equal keys in Gaussian order.
    order = np.lexsort((np.arange(cloud.count), r.keys))
    assert np.array_equal(r.values, order) or np.array_equal(r.keys, np.sort(r.keys))
    assert (np.diff(r.keys.astype(np.int64)) >= 0).all()
    same = np.diff(r.keys.astype(np.int64)) == 0
    assert (np.diff(r.values.astype(np.int64))[same] > 0).all()
    # Entries: sorted by tile; within a tile in depth order (depth rank ascending).
    assert (np.diff(r.entry_keys.astype(np.int64)) >= 0).all()
    rank = np.empty(cloud.count, np.int64)
    rank[r.values] = np.arange(cloud.count)
    tiles_equal = np.diff(r.entry_keys.astype(np.int64)) == 0
    assert (np.diff(rank[r.entry_values])[tiles_equal] > 0).all()
    # Ranges partition [0, E): each tile's run holds exactly its entries.
    total = 0
    for t in range(m.tiles):
        a, b = r.ranges[t]
        if b > a:
            assert (r.entry_keys[a:b] == t).all()
            total += b - a
    assert total == r.entries


def test_degenerate_gaussians_do_not_change_the_picture():
    """test_train.cpp:544-586: NaN position, overflowing scale, short quaternion,
    NaN opacity and NaN SH leave the render bitwise unchanged."""
    cam = F.train_camera(72, 56)
    healthy = F.random_cloud(40, 1, 41)
    base = O.Model(healthy, cam).forward().image
    c = F.Cloud.zeros(45, 1)
    for k in ("positions", "scales", "rotations", "opacities", "sh"):
        getattr(c, k)[:40] = getattr(healthy, k)
    for i in range(40, 45):
        c.positions[i, 2] = 4.0
        c.scales[i] = math.log(0.3)
        c.rotations[i, 0] = 1.0
        c.opacities[i] = 2.0
    with np.errstate(invalid="ignore"):
        c.positions[40, 0] = np.nan
        c.scales[41, 0] = 200.0
        c.rotations[42, 0] = 1e-5
        c.opacities[43] = np.nan
        c.sh[44, 0, 0] = np.nan
        image = O.Model(c, cam).forward().image
    assert np.array_equal(image.view(np.uint32), base.view(np.uint32))


def test_fixed_point_sums_equal_float_sums_within_rounding():
    cloud, cam, target = _case("ref300")
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2)
    fl = m.backward_float(r).raw.copy()
    m.backward_fixed(r)
    scale = m.fixed_scale()
    waves = r.entries * 4  # at most one quantised total per (entry, wave)
    assert np.abs(r.raw_fixed / scale - fl).max() <= 0.5 * waves / scale + 1e-6


def test_first_adam_step_moves_by_the_rate():
    """test_train.cpp:450-491 on the model: m / (1 - b1) = g and sqrt(v / (1 - b2)) = |g|
    in the first step, so each parameter moves by its rate against its gradient's
    sign (to the rounding of bf16 moments, which Adam sees only after the step)."""
    import ctypes as C
    cloud, cam, target = F.random_cloud(8, 1, 5), F.train_camera(40, 32), F.smooth_target(40, 32, 2)
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_float(r)
    g, _ = m.gradients(r)
    n = cloud.count
    pos = cloud.positions.copy().reshape(-1); sc = cloud.scales.copy().reshape(-1)
    rot = cloud.rotations.copy().reshape(-1); op = cloud.opacities.copy()
    words = m.words.copy()
    states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 12)]
    stats = np.zeros((n, 8), np.float32)
    raw = r.raw.copy()
    rates = (1.6e-4, 0.005, 0.001, 0.05, 0.02)
    opt = m.chain_options(mode=2, flags=4, rates=rates, step=1)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    assert O.lib().ob_chain(p(pos), p(sc), p(rot), p(op), p(words), None, C.byref(m.camera), p(raw), p(r.splats), p(stats),
                            *[p(s) for s in states], None, None, None, None, None, C.byref(opt)) == 0
    assert O.lib().ob_sh_step(p(words), p(states[4]), p(raw), C.byref(m.camera), C.byref(opt)) == 0
    moved = op - cloud.opacities
    nz = g[3] != 0
    assert nz.any()
    assert np.allclose(moved[nz], -np.sign(g[3][nz]) * 0.05, atol=2e-5)
    sh = F.unpack_half_sh(words, 4).reshape(n, 12)
    before = cloud.sh.reshape(n, 12)
    gsh = g[4].reshape(n, 12)
    for i in range(n):
        for j in range(12):
            if gsh[i, j] == 0:
                continue
            rate = 0.02 if j < 3 else 0.02 * 0.05
            expected = before[i, j] - math.copysign(rate, gsh[i, j])
            tol = 2e-5 if j < 3 else max(2e-5, abs(expected) / 512)
            assert abs(sh[i, j] - expected) <= tol


def _bf16_neighbours(x):
    """The bf16 values either side of each float32 (truncation and one bf16 ulp
    away from zero), as float32."""
    bits = np.asarray(x, np.float32).view(np.uint32)
    low = (bits & 0xFFFF0000).view(np.float32)
    high = ((bits & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32)
    return low, high


def test_adam_moments_round_to_bf16_as_stated():
    """train_step.glsl:49-60 and bf16.glsl:10-23 on the model: after the first step
    the first moment is (1 - b1) g rounded to the nearest bf16 (ties to even), and
    the second moment is ((1 - b2) g) g rounded to one of its two bf16 neighbours
    (stochastic rounding). Checked by distance in float64, not by the bit formula."""
    import ctypes as C
    cloud, cam, target = F.random_cloud(8, 1, 5), F.train_camera(40, 32), F.smooth_target(40, 32, 2)
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_float(r)
    g, _ = m.gradients(r)
    n = cloud.count
    pos = cloud.positions.copy().reshape(-1); sc = cloud.scales.copy().reshape(-1)
    rot = cloud.rotations.copy().reshape(-1); op = cloud.opacities.copy()
    words = m.words.copy()
    states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 12)]
    stats = np.zeros((n, 8), np.float32)
    opt = m.chain_options(mode=2, flags=4, rates=(1.6e-4, 0.005, 0.001, 0.05, 0.02), step=1)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    assert O.lib().ob_chain(p(pos), p(sc), p(rot), p(op), p(words), None, C.byref(m.camera), p(r.raw.copy()), p(r.splats),
                            p(stats), *[p(s) for s in states], None, None, None, None, None, C.byref(opt)) == 0
    one = np.float32(1)
    checked = 0
    for grad, state in zip(g[:4], states[:4]):
        grad = np.asarray(grad, np.float32).reshape(-1)
        touched = grad != 0
        m1 = ((one - np.float32(0.9)) * grad)[touched]
        v1 = (((one - np.float32(0.999)) * grad) * grad)[touched]
        got_m = ((state[touched] & 0xFFFF) << 16).astype(np.uint32).view(np.float32)
        got_v = (state[touched] & 0xFFFF0000).astype(np.uint32).view(np.float32)
        lo, hi = _bf16_neighbours(m1)
        dlo, dhi = np.abs(lo.astype(np.float64) - m1), np.abs(hi.astype(np.float64) - m1)
        even_lo = ((lo.view(np.uint32) >> 16) & 1) == 0
        nearest = np.where((dlo < dhi) | ((dlo == dhi) & even_lo), lo, hi)
        assert np.array_equal(got_m.view(np.uint32), nearest.view(np.uint32))
        lo, hi = _bf16_neighbours(v1)
        assert np.all((got_v == lo) | (got_v == hi))
        assert np.all(state[~touched] == 0)  # selective Adam: untouched state stays zero
        checked += int(touched.sum())
    assert checked > 20


def test_adam_from_gradients_equals_the_chain_rules_step():
    """The batched step's Adam (ob_adam, from gradient arrays) and the chain
    rule's own step (ob_chain in step mode with ob_sh_step) move a view's
    parameters identically: same rule, same rounding stream."""
    cloud, cam, target = _case("ref60")
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_fixed(r)
    g, _ = m.gradients(r)
    rates = (1.6e-4, 0.005, 0.001, 0.05, 0.025)
    want = m.step(r, rates)
    got = O.adam_from_gradients(cloud, m.words, g, rates)
    for a, b in zip(got, want):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    assert not np.array_equal(got[3], cloud.opacities)


@pytest.mark.parametrize("fixed", [False, True])
def test_bf16_sh_model_matches_fp64_reference(fixed):
    """SH stored as bf16: Oracle B against Oracle A run on the same bf16-rounded
    coefficients, with the reference's tolerances; the arithmetic stays fp32."""
    cloud, cam, target = _case("deg3")
    stored = cloud.copy()
    stored.sh = F.unpack_bf16_sh(F.pack_bf16_sh(cloud.sh), cloud.coefficients).reshape(cloud.sh.shape)
    assert not np.array_equal(stored.sh, cloud.sh)  # the rounding did something
    ref = O.render(stored, cam, threads=2)
    m = O.Model(cloud, cam, sh_storage=1)
    r = m.forward()
    assert np.abs(r.image[:, :3] - ref.rgb).max() < 1e-3
    ref_loss = O.loss(ref, target, 0.2, threads=2)
    m.loss(r, target, 0.2)
    ref_g = O.gradients(stored, cam, ref_loss.adjoint, threads=2)
    (m.backward_fixed if fixed else m.backward_float)(r)
    g, _ = m.gradients(r)
    cos, ratio = (0.97, 0.1) if fixed else (0.999, 0.02)
    for got, want, what in zip(g, (ref_g.positions, ref_g.scales, ref_g.rotations, ref_g.opacities, ref_g.sh),
                               ("positions", "scales", "rotations", "opacities", "sh")):
        check_close(got, want, what, cos, ratio)


def test_bf16_sh_step_rounds_to_a_neighbour():
    """After a step with SH stored as bf16, every coefficient is one of the two
    bf16 values around the fp32 result (stochastic rounding), and the DC terms
    moved by about the rate like every other parameter."""
    cloud, cam, target = _case("deg3")
    rates = (1.6e-4, 0.005, 0.001, 0.05, 0.025)
    stepped = []
    for storage in (0, 1):
        m = O.Model(cloud, cam, sh_storage=storage)
        r = m.forward()
        m.loss(r, target, 0.2)
        m.backward_fixed(r)
        stepped.append(m.step(r, rates, rest_scale=0.05, step=1))
    words = stepped[1][4]
    got = F.unpack_bf16_sh(words, cloud.coefficients).reshape(cloud.count, -1).astype(np.float64)
    before = F.unpack_bf16_sh(F.pack_bf16_sh(cloud.sh), cloud.coefficients).reshape(cloud.count, -1).astype(np.float64)
    moved = got != before
    assert moved[:, :3].mean() > 0.5
    # The two bf16 neighbours of any float x bracket x: truncation and one ulp away.
    dc = np.abs(got[:, :3] - before[:, :3])
    assert dc[moved[:, :3]].max() < 0.025 * 1.02 + np.abs(before[:, :3]).max() * 2 ** -7
    # Geometry does not depend on the storage beyond the colour: positions move alike in sign.
    assert np.mean(np.sign(stepped[0][0] - cloud.positions.reshape(-1)) ==
                   np.sign(stepped[1][0] - cloud.positions.reshape(-1))) > 0.95


def _needle_scene(length=20.0, thin=0.03):
    """30 ordinary Gaussians and three long, thin ones lying across the image,
    whose 2D covariance is nearly singular: a c - b b cancels in float."""
    base = F.random_cloud(30, 1, 29)
    c = F.Cloud.zeros(33, 1)
    for k in ("positions", "scales", "rotations", "opacities", "sh"):
        getattr(c, k)[:30] = getattr(base, k)
    for j, angle in enumerate([45.0, -30.0, 60.0]):
        i = 30 + j
        c.positions[i] = (0.1 * j, -0.1 * j, 4.0 + j)
        c.scales[i] = (math.log(length), math.log(thin), math.log(thin))
        h = math.radians(angle) / 2
        c.rotations[i] = (math.cos(h), 0.0, 0.0, math.sin(h))
        c.opacities[i] = 2.0
        c.sh[i, 0] = (0.8, -0.5, 0.3)
    return c


@pytest.mark.parametrize("length,thin", [(20.0, 0.03), (60.0, 0.03), (200.0, 0.05), (600.0, 0.3)])
def test_near_degenerate_covariance_in_double_matches_fp64(length, thin):
    """Long, thin splats lying across the image: in float their small axis is
    lost in the covariance sums (image errors of 2e-3 to 0.2 against FP64 when
    only float is used); built and inverted in double they render within a
    tenth of the reference's forward tolerance."""
    cloud, cam = _needle_scene(length, thin), F.train_camera(72, 56)
    r = O.Model(cloud, cam).forward()
    ref = O.render(cloud, cam, threads=2)
    assert np.abs(r.image[:, :3] - ref.rgb).max() < 1e-4
    assert np.abs(r.image[:, 3] - ref.log_t).max() < 1e-3
    for i in (30, 31, 32):
        qx, qy, qz = r.splats[i, :3].astype(np.float64)
        eig = np.linalg.eigvalsh(np.array([[qx, qy], [qy, qz]]))
        assert eig[0] > 0 and eig[1] / eig[0] > 1e3  # the splat is a needle
