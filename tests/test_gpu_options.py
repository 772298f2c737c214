"""The reference's option paths on the GPU, each bit for bit against Oracle B
(whose branches tests/test_oracle_b_options.py pins against the FP64 Oracle A)
and within the reference's bars against Oracle A:

* the 2D Mip filter and the 3D filter (TrainerOptions::mipFilter2d,
  Trainer::set_filter3d; tests/test_train.cpp:184-216);
* a background colour (TrainerOptions::background);
* the opacity and scale penalties (tests/test_train.cpp:295-321);
* set_active_sh_degree below the cloud's degree (trainer.hpp:151);
* SH dropout in a step (Trainer::set_sh_dropout; tests/test_train.cpp:659-718);
* non-selective Adam (TrainerOptions::selectiveAdam = false, trainer.cpp:455-456);
* all of them at once in a step, at 20k Gaussians and 1280x720.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

RATES = vsplat.LearningRates(position=1.6e-4, scale=0.005, rotation=0.001, opacity=0.05, sh=0.025, sh_rest_scale=0.05)
RATE_TUPLE = (RATES.position, RATES.scale, RATES.rotation, RATES.opacity, RATES.sh)
GROUPS = ("positions", "scales", "rotations", "opacities", "sh")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1)


def _check_close(a, b, what, min_cos, max_ratio):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    cos = (a @ b) / np.sqrt((a @ a) * (b @ b))
    ratio = np.sqrt((a @ a) / (b @ b))
    assert cos > min_cos, f"{what}: cosine {cos}"
    assert abs(ratio - 1) < max_ratio, f"{what}: ratio {ratio}"


def _filter(n):
    return (0.001 * (np.arange(n) % 7)).astype(np.float32)


OPTIONS = {
    "mip": dict(options=dict(mip_filter_2d=True), model=dict(mip_filter_2d=True), ref=dict(mip_filter_2d=True)),
    "filter3d": dict(filter3d=True, options={}, model={}, ref={}),
    "mip+filter3d": dict(filter3d=True, options=dict(mip_filter_2d=True), model=dict(mip_filter_2d=True),
                         ref=dict(mip_filter_2d=True)),
    "background": dict(options=dict(background=(0.2, 0.5, 0.8)), model=dict(background=(0.2, 0.5, 0.8)),
                       ref=dict(background=(0.2, 0.5, 0.8))),
    "penalties": dict(options=dict(opacity_reg=0.5, scale_reg=0.25), model={}, ref=dict(opacity_reg=0.5, scale_reg=0.25),
                      penalties=(0.5, 0.25)),
}


def _setup(name, cloud):
    spec = OPTIONS[name]
    f = _filter(cloud.count) if spec.get("filter3d") else None
    model_kw = dict(spec["model"])
    ref_kw = dict(spec["ref"])
    if f is not None:
        model_kw["filter3d"] = f
        ref_kw["filter3d"] = f
    return spec, f, model_kw, ref_kw


@pytest.mark.parametrize("name", list(OPTIONS))
def test_option_gradients_bit_exact(name):
    cloud, cam, target = F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    spec, f, model_kw, _ = _setup(name, cloud)
    reg = spec.get("penalties", (0.0, 0.0))
    m = O.Model(cloud, cam, **model_kw)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_fixed(r)
    g_model, st_model = m.gradients(r, opacity_reg=reg[0], scale_reg=reg[1])
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, **spec["options"])) as t:
        if f is not None:
            t.set_filter3d(f)
        stats = t.compute_gradients(cam, target)
        assert stats.visible_splats == r.visible and stats.tile_entries == r.entries
        visible = r.values[r.keys != 0xFFFFFFFF]
        np.testing.assert_array_equal(_bits(t.inspect("splats")[visible]), _bits(r.splats[visible]))
        np.testing.assert_array_equal(_bits(t.inspect("image")), _bits(r.image))
        np.testing.assert_array_equal(t.inspect("raw"), r.raw_fixed)
        g = t.gradients()
        for got, want, what in zip((g.positions, g.scales, g.rotations, g.opacities, g.sh), g_model, GROUPS):
            np.testing.assert_array_equal(_bits(got), _bits(want), err_msg=what)
        np.testing.assert_array_equal(_bits(t.densify_stats()), _bits(st_model))


@pytest.mark.parametrize("fixed", [False, True])
@pytest.mark.parametrize("name", list(OPTIONS))
def test_option_against_fp64(name, fixed):
    # tests/test_train.cpp:184-216 (Mip), with the reference's bars.
    cloud, cam, target = F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.random_target(72, 56, 11)
    spec, f, _, ref_kw = _setup(name, cloud)
    opts = O.RefOptions(**ref_kw)
    ref = O.render(cloud, cam, opts)
    ref_loss = O.loss(ref, target, 0.2)
    ref_g = O.gradients(cloud, cam, ref_loss.adjoint, opts)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=fixed, **spec["options"])) as t:
        if f is not None:
            t.set_filter3d(f)
        image = t.render(cam)
        assert np.abs(image[:, :3] - ref.rgb).max() < 1e-3
        assert np.abs(image[:, 3] - ref.log_t).max() < 1e-2
        stats = t.compute_gradients(cam, target)
        assert abs(stats.loss - ref_loss.loss) < 1e-3
        g = t.gradients()
    cos, ratio = (0.97, 0.1) if fixed else (0.999, 0.02)
    for got, want, what in zip((g.positions, g.scales, g.rotations, g.opacities, g.sh),
                               (ref_g.positions, ref_g.scales, ref_g.rotations, ref_g.opacities, ref_g.sh), GROUPS):
        _check_close(got, want, what, cos, ratio)


def test_penalties_behind_the_camera_are_the_closed_form():
    # tests/test_train.cpp:295-321
    cam, target = F.train_camera(64, 48), F.random_target(64, 48, 5)
    cloud = F.random_cloud(20, 1, 9)
    cloud.positions[:, 2] = -5.0
    n = cloud.count
    for reg in (0.0, 0.5):
        with vsplat.Trainer(cloud, vsplat.TrainerOptions(opacity_reg=reg, scale_reg=reg * 0.5)) as t:
            t.compute_gradients(cam, target)
            g = t.gradients()
        o = 1.0 / (1.0 + np.exp(-cloud.opacities.astype(np.float64)))
        np.testing.assert_allclose(g.opacities, reg / n * o * (1 - o), atol=1e-6)
        np.testing.assert_allclose(g.scales.reshape(-1, 3), reg * 0.5 / n * np.exp(cloud.scales.astype(np.float64)), atol=1e-6)
        assert np.all(g.positions[:, 0] == 0)


@pytest.mark.parametrize("active", [0, 1, 2])
def test_active_degree_below_the_cloud_bit_exact(active):
    cloud, cam, target = F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    m = O.Model(cloud, cam, active_degree=active)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_fixed(r)
    g_model, _ = m.gradients(r)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        t.set_active_sh_degree(active)
        assert t.active_sh_degree() == active
        t.compute_gradients(cam, target)
        np.testing.assert_array_equal(_bits(t.inspect("image")), _bits(r.image))
        g = t.gradients()
        for got, want, what in zip((g.positions, g.scales, g.rotations, g.opacities, g.sh), g_model, GROUPS):
            np.testing.assert_array_equal(_bits(got), _bits(want), err_msg=what)
    assert np.all(g.sh[:, (active + 1) ** 2:] == 0)


def _model_step(cloud, cam, target, *, model_kw=None, selective=True, reg=(0.0, 0.0), sh_storage=0):
    m = O.Model(cloud, cam, sh_storage=sh_storage, **(model_kw or {}))
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_fixed(r)
    return m.step(r, RATE_TUPLE, rest_scale=RATES.sh_rest_scale, step=1, selective=selective, opacity_reg=reg[0],
                  scale_reg=reg[1])


def _assert_cloud(c, want, coefficients, sh_storage=0):
    pos, sc, rot, op, words = want
    np.testing.assert_array_equal(_bits(c.positions), _bits(pos))
    np.testing.assert_array_equal(_bits(c.scales), _bits(sc))
    np.testing.assert_array_equal(_bits(c.rotations), _bits(rot))
    np.testing.assert_array_equal(_bits(c.opacities), _bits(op))
    np.testing.assert_array_equal(_bits(c.sh), _bits(F.unpack_sh(words, coefficients, sh_storage)))


@pytest.mark.parametrize("retained,sh_storage", [(0, 0), (1, 0), (1, 1)])
def test_sh_dropout_step_bit_exact(retained, sh_storage):
    cloud, cam, target = F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    dropout = (0.5, retained, 9)
    want = _model_step(cloud, cam, target, model_kw=dict(sh_dropout=dropout), sh_storage=sh_storage)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, sh_storage=sh_storage)) as t:
        t.set_sh_dropout(*dropout)
        t.step(cam, target, RATES)
        _assert_cloud(t.cloud(), want, 16, sh_storage)


def test_sh_dropout_trains_the_retained_degrees_only():
    # tests/test_train.cpp:659-718
    cam = F.train_camera(72, 56)
    target = F.random_target(72, 56, 17)
    cloud = F.random_cloud(300, 3, 7)
    full = O.render(cloud, cam)
    rates = vsplat.LearningRates(position=0.0, scale=0.0, rotation=0.0, opacity=0.0, sh=0.01)

    def higher_nonzero(sh):
        return int(np.count_nonzero(np.any(sh.reshape(cloud.count, 16, 3)[:, 1:] != 0, axis=(1, 2))))

    with vsplat.Trainer(cloud) as t:
        t.set_active_sh_degree(3)

        def trained():
            before = t.cloud().sh
            t.step(cam, target, rates)
            return higher_nonzero(t.cloud().sh - before)

        t.set_sh_dropout(1.0, 0, 1)
        assert np.abs(t.render(cam)[:, :3] - full.rgb).max() < 1e-3
        assert trained() == 0
        t.compute_gradients(cam, target)  # statistics: not training, not dropped
        assert higher_nonzero(t.gradients().sh) > 250
        t.set_sh_dropout(0.5, 0, 9)
        kept = trained()
        print(f"SH dropout at 0.5: {kept} of {cloud.count} kept their higher degrees")
        assert 100 < kept < 200
        t.set_sh_dropout(0.0, 0, 1)
        assert trained() > 250


STATE = ("state_positions", "state_scales", "state_rotations", "state_opacities", "state_sh")


@pytest.mark.parametrize("sh_storage", [0, 1])
def test_non_selective_adam_bit_exact(sh_storage):
    """Two steps: the first sees every Gaussian, the second only the right half
    of the view, so that Gaussians with moments get no gradient. Non-selective
    Adam moves them by their momentum; selective Adam leaves them bit for bit
    (trainer.cpp:455-456). Each step against Oracle B from the GPU's state."""
    cloud, cam, target = F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    half = dataclasses.replace(cam, width=36, cx=cam.cx - 36.0)
    target2 = F.smooth_target(36, 56, 3)
    for selective in (False, True):
        options = vsplat.TrainerOptions(force_fixed_point=True, selective_adam=selective, sh_storage=sh_storage)
        with vsplat.Trainer(cloud, options) as t:
            t.step(cam, target, RATES)
            first = t.cloud()
            _assert_cloud(first, _model_step(cloud, cam, target, selective=selective, sh_storage=sh_storage), 16, sh_storage)
            states = [t.inspect(k).reshape(-1).copy() for k in STATE]
            t.step(half, target2, RATES)
            second = t.cloud()
            m = O.Model(first, half, sh_storage=sh_storage)
            r = m.forward()
            m.loss(r, target2, 0.2)
            m.backward_fixed(r)
            raw = r.raw.copy()
            want = m.step(r, RATE_TUPLE, rest_scale=RATES.sh_rest_scale, step=2, selective=selective, states=states)
            _assert_cloud(second, want, 16, sh_storage)
        unseen = ~np.any(raw != 0, axis=1)
        assert unseen.sum() > 20
        moved = np.any(_bits(second.positions).reshape(-1, 3) != _bits(first.positions).reshape(-1, 3), axis=1)
        if selective:
            assert not moved[unseen].any()
        else:
            assert moved[unseen].sum() > 20


def test_every_option_at_once_in_a_step_at_full_size():
    """Mip, the 3D filter, a background, the penalties, dropout and
    non-selective Adam together, at config 1's distributions (20k Gaussians,
    1280x720): the step bit for bit."""
    w = F.config(1, count=20_000)
    cloud, cam, target = w.cloud, w.cameras[0], w.targets[0]
    f = _filter(cloud.count)
    bg = (0.1, 0.2, 0.3)
    dropout = (0.3, 1, 5)
    want = _model_step(cloud, cam, target, model_kw=dict(mip_filter_2d=True, filter3d=f, background=bg, sh_dropout=dropout),
                       selective=False, reg=(0.5, 0.25))
    options = vsplat.TrainerOptions(force_fixed_point=True, mip_filter_2d=True, background=bg, opacity_reg=0.5, scale_reg=0.25,
                                    selective_adam=False)
    with vsplat.Trainer(cloud, options) as t:
        t.set_filter3d(f)
        t.set_sh_dropout(*dropout)
        t.step(cam, target, RATES)
        _assert_cloud(t.cloud(), want, 16)
