"""Exposure and view scoring on the GPU (Trainer::set_exposure, Trainer::score_view;
trainer.cpp:397-425, 519-532):

* score_view: the densification statistics it leaves equal Oracle B's scored
  pass bit for bit (fixed point); the marked pixels and marked loss equal
  Oracle A's FP64 count within a pixel either way as the reference's test
  allows (tests/test_train.cpp:588-621), with fixed-point and float sums; an
  ordinary iteration afterwards adds nothing to them; the edge score is
  positive on a target's edge and zero away from it (tests/test_train.cpp:623-657);
* set_exposure: the loss, the fixed-point sums, the gradients and the
  exposure gradient equal Oracle B's bit for bit; the neutral exposure changes
  nothing; the queued and batched calls refuse a non-neutral one.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

EXPOSURE = np.array([1.1, 0.05, -0.02, 0.03, 0.02, 0.9, 0.04, -0.01, -0.03, 0.06, 1.05, 0.02], np.float32)
GROUPS = ("positions", "scales", "rotations", "opacities", "sh")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1)


def _score_case():
    return F.random_cloud(60, 1, 47), F.train_camera(72, 56), F.random_target(72, 56, 43)


def test_score_view_matches_the_model_bit_for_bit():
    cloud, cam, target = _score_case()
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        stats = t.score_view(cam, target, 0.5, 2.0)
        got = t.densify_stats()
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2, score_share=0.5)
    m.backward_fixed(r)
    want = O.score_statistics(m.statistics(r), r.raw, 3.0 * cam.width * cam.height * 2.0, stats.loss)
    print(f"marked pixels {got[:, 5].sum():.1f}, edge score {got[:, 7].sum():.4f}", flush=True)
    assert got[:, 5].sum() > 0
    np.testing.assert_array_equal(_bits(got), _bits(want))


@pytest.mark.parametrize("fixed", [True, False])
def test_score_view_counts_the_high_error_pixels(fixed):
    cloud, cam, target = _score_case()
    cpu = O.marked_pixels(cloud, cam, target, 0.5)
    assert cpu.sum() > 0
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=fixed)) as t:
        t.compute_gradients(cam, target)  # an ordinary iteration counts nothing
        assert not np.any(t.densify_stats()[:, 5:])
        stats = t.score_view(cam, target, 0.5, 2.0)
        scored = t.densify_stats()
        t.compute_gradients(cam, target)
        after = t.densify_stats()
    np.testing.assert_allclose(scored[:, 5], 2.0 * cpu, atol=3.0)
    np.testing.assert_allclose(scored[:, 6], 2.0 * cpu * stats.loss, atol=3.0)
    np.testing.assert_array_equal(after[:, 5:], scored[:, 5:])


def test_edge_score_on_and_off_the_edge():
    w, h = 96, 64
    cam = F.train_camera(w, h)
    target = np.full((h, w, 4), 255, np.uint8)
    target[:, : w // 2, :3] = 40
    target[:, w // 2:, :3] = 220
    cloud = F.Cloud.zeros(16, 0)
    for i in range(16):
        cloud.positions[i] = [0.0 if i < 8 else 0.8, (i % 8 - 3.5) * 0.1, 2.0]
        cloud.scales[i] = np.log(0.02)
        cloud.rotations[i, 0] = 1.0
    with vsplat.Trainer(cloud, vsplat.TrainerOptions()) as t:
        t.score_view(cam, target, 0.5, 1.0)
        s = t.densify_stats()
    print(f"edge score on the edge {s[:8, 7].mean():.4f}, away {s[8:, 7].mean():.4f}", flush=True)
    assert s[:8, 7].sum() > 0.0
    assert s[8:, 7].sum() == 0.0
    assert np.all(s[:, 5] > 0.0)


@pytest.mark.parametrize("name", ["ref60", "deg3"])
def test_exposure_matches_the_model_bit_for_bit(name):
    if name == "ref60":
        cloud, cam, target = F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.smooth_target(72, 56, 11)
    else:
        cloud, cam, target = F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        t.set_exposure(EXPOSURE)
        stats = t.compute_gradients(cam, target)
        g = t.gradients()
        raw = t.inspect("raw")
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2, exposure=EXPOSURE)
    m.backward_fixed(r)
    want, _ = m.gradients(r)
    np.testing.assert_array_equal(raw.reshape(-1), r.raw_fixed.reshape(-1))
    for name_, w in zip(GROUPS, want):
        np.testing.assert_array_equal(_bits(getattr(g, name_)), _bits(w))
    np.testing.assert_array_equal(_bits(np.array(stats.exposure_gradient, np.float32)), _bits(r.exposure_gradient))
    assert np.any(np.array(stats.exposure_gradient) != 0.0)


def test_neutral_exposure_changes_nothing_and_queues_refuse_others():
    cloud, cam, target = F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.smooth_target(72, 56, 11)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as a, \
         vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as b:
        b.set_exposure(O.NEUTRAL_EXPOSURE)
        sa, sb = a.compute_gradients(cam, target), b.compute_gradients(cam, target)
        ga, gb = a.gradients(), b.gradients()
        assert sa.loss == sb.loss and not any(sb.exposure_gradient)
        for name in GROUPS:
            np.testing.assert_array_equal(_bits(getattr(ga, name)), _bits(getattr(gb, name)))
        b.set_exposure(EXPOSURE)
        with vsplat.DeviceBuffer(target) as d:
            with pytest.raises(vsplat.Unsupported):
                b.enqueue(cam, d.ptr, vsplat.MODE_GRADIENTS)
        with pytest.raises(vsplat.Unsupported):
            b.step_views([cam], [target])
        b.set_exposure(O.NEUTRAL_EXPOSURE)
        b.step_views([cam], [target])
