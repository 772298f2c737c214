"""Oracle B's option branches against Oracle A, whose gradients are pinned by
FP64 finite differences (tests/test_oracle_a.py, including the Mip filters),
with the reference's own tolerances (tests/test_train.cpp:38-53, 184-216,
295-321, 659-718): image <= 1e-3, log T <= 1e-2, loss <= 1e-3, gradient
cosine > 0.999 and length within 2 % with float sums, > 0.97 and 10 % with
fixed point.

* the 2D Mip filter and the 3D smoothing filter, apart and together
  (train_preprocess.comp Mip/filter lines; train_gradient.comp k2/k3);
* a background colour (train_forward.comp: rgb + T bg, and its derivative);
* the opacity and scale penalties: closed form where no pixel sees the
  Gaussians, and with the image's gradient where they are seen;
* an active SH degree below the cloud's: Oracle A on the cloud cut to that
  degree, and exactly zero gradients above it;
* SH dropout (train_types.glsl:165-170): the dropped Gaussians are those whose
  hash draw is under the probability (the hash is pinned in
  tests/test_rounding.py); Oracle A on the cloud with their coefficients above
  the retained degree set to zero, and exactly zero gradients for those.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F

GROUPS = ("positions", "scales", "rotations", "opacities", "sh")


def check_close(a, b, what, min_cos, max_ratio):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    cos = (a @ b) / math.sqrt((a @ a) * (b @ b))
    ratio = math.sqrt((a @ a) / (b @ b))
    assert cos > min_cos, f"{what}: cosine {cos}"
    assert abs(ratio - 1) < max_ratio, f"{what}: ratio {ratio}"


def _compare(cloud, cam, target, model, ref_options, fixed, *, ref_cloud=None, opacity_reg=0.0, scale_reg=0.0,
             groups=GROUPS):
    """Model forward/loss/backward/chain against Oracle A on ref_cloud (default cloud)."""
    ref_cloud = cloud if ref_cloud is None else ref_cloud
    ref = O.render(ref_cloud, cam, ref_options, threads=2)
    r = model.forward()
    assert np.abs(r.image[:, :3] - ref.rgb).max() < 1e-3
    assert np.abs(r.image[:, 3] - ref.log_t).max() < 1e-2
    ref_loss = O.loss(ref, target, 0.2, threads=2)
    model.loss(r, target, 0.2)
    t = r.tile_totals.astype(np.float64).sum(0) / (3 * cam.width * cam.height)
    assert abs(t[1] - ref_loss.l1) < 1e-3 and abs(t[0] - ref_loss.ssim) < 1e-3
    ref_g = O.gradients(ref_cloud, cam, ref_loss.adjoint, ref_options, threads=2)
    (model.backward_fixed if fixed else model.backward_float)(r)
    g, _ = model.gradients(r, opacity_reg=opacity_reg, scale_reg=scale_reg)
    cos, ratio = (0.97, 0.1) if fixed else (0.999, 0.02)
    want = dict(zip(GROUPS, (ref_g.positions, ref_g.scales, ref_g.rotations, ref_g.opacities, ref_g.sh)))
    for got, name in zip(g, GROUPS):
        if name in groups:
            check_close(got, want[name], name, cos, ratio)
    return g, ref_g


@pytest.mark.parametrize("fixed", [False, True])
@pytest.mark.parametrize("mip,filter3d", [(True, False), (False, True), (True, True)])
def test_mip_and_3d_filter(mip, filter3d, fixed):
    # tests/test_train.cpp:184-216
    cloud, cam, target = F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.random_target(72, 56, 11)
    f = (0.001 * (np.arange(cloud.count) % 7)).astype(np.float32) if filter3d else None
    model = O.Model(cloud, cam, mip_filter_2d=mip, filter3d=f)
    _compare(cloud, cam, target, model, O.RefOptions(mip_filter_2d=mip, filter3d=f), fixed)


@pytest.mark.parametrize("fixed", [False, True])
def test_background(fixed):
    bg = (0.2, 0.5, 0.8)
    cloud, cam, target = F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.smooth_target(72, 56, 11)
    model = O.Model(cloud, cam, background=bg)
    _compare(cloud, cam, target, model, O.RefOptions(background=bg), fixed)


def test_penalties_alone_are_the_closed_form():
    # tests/test_train.cpp:295-321: behind the camera, the penalties are the whole gradient.
    cam, target = F.train_camera(64, 48), F.random_target(64, 48, 5)
    cloud = F.random_cloud(20, 1, 9)
    cloud.positions[:, 2] = -5.0
    n = cloud.count
    for reg in (0.0, 0.5):
        model = O.Model(cloud, cam)
        r = model.forward()
        model.loss(r, target, 0.2)
        model.backward_fixed(r)
        g, _ = model.gradients(r, opacity_reg=reg, scale_reg=reg * 0.5)
        o = 1.0 / (1.0 + np.exp(-cloud.opacities.astype(np.float64)))
        np.testing.assert_allclose(g[3], reg / n * o * (1 - o), atol=1e-6)
        np.testing.assert_allclose(g[1].reshape(-1, 3), reg * 0.5 / n * np.exp(cloud.scales.astype(np.float64)), atol=1e-6)
        assert np.all(g[0] == 0)


@pytest.mark.parametrize("fixed", [False, True])
def test_penalties_with_the_image(fixed):
    cloud, cam, target = F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.smooth_target(72, 56, 11)
    model = O.Model(cloud, cam)
    _compare(cloud, cam, target, model, O.RefOptions(opacity_reg=0.5, scale_reg=0.25), fixed, opacity_reg=0.5,
             scale_reg=0.25)


def _cut(cloud, degree):
    c = cloud.copy()
    c.sh = np.ascontiguousarray(cloud.sh[:, : (degree + 1) ** 2])
    c.sh_degree = degree
    return c


@pytest.mark.parametrize("active", [0, 1, 2])
def test_active_degree_below_the_cloud(active):
    cloud, cam, target = F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    model = O.Model(cloud, cam, active_degree=active)
    g, ref_g = _compare(cloud, cam, target, model, None, False, ref_cloud=_cut(cloud, active), groups=GROUPS[:4])
    sh = g[4].reshape(cloud.count, 16, 3)
    used = (active + 1) ** 2
    assert np.all(sh[:, used:] == 0)
    check_close(sh[:, :used], ref_g.sh, "sh", 0.999, 0.02)


@pytest.mark.parametrize("retained", [0, 1])
def test_sh_dropout_equals_the_cut_coefficients(retained):
    p, seed = 0.5, 9
    cloud, cam, target = F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    dropped = O.train_uniform(np.arange(cloud.count, dtype=np.uint32), seed) < np.float32(p)
    assert 100 < np.count_nonzero(dropped) < 200  # the reference's own bounds (test_train.cpp:706-708)
    keep = (retained + 1) ** 2
    ref_cloud = cloud.copy()
    ref_cloud.sh[dropped, keep:] = 0.0
    model = O.Model(cloud, cam, sh_dropout=(p, retained, seed))
    g, ref_g = _compare(cloud, cam, target, model, None, False, ref_cloud=ref_cloud, groups=GROUPS[:4])
    sh = g[4].reshape(cloud.count, 16, 3)
    assert np.all(sh[dropped, keep:] == 0)
    assert np.count_nonzero(sh[~dropped, keep:]) > 0
    # The coefficients a dropped Gaussian does not use get no gradient; Oracle A
    # differentiates the cut cloud's zeros too, so those are left out of it.
    want = ref_g.sh.reshape(cloud.count, 16, 3).copy()
    want[dropped, keep:] = 0.0
    check_close(sh, want, "sh", 0.999, 0.02)
