"""The pose gradient on the GPU (TrainerOptions::poseGradients,
StepStats::poseGradient; trainer.cpp:455, 630-637, train_gradient.comp:276-337):

* compute_gradients fills it equal to Oracle B's bit for bit (fixed point;
  the same group tree and ordered double sum), and leaves every gradient as
  without it;
* a step fills the same pose gradient (taken before the step changes
  anything); without the option, and from a batch of views, it stays zero;
* it agrees with central differences of Oracle A's loss under refined_pose as
  the reference's test requires (tests/test_train.cpp:722-742).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

GROUPS = ("positions", "scales", "rotations", "opacities", "sh")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1)


def _case(count=300, degree=0, seed=59):
    return F.random_cloud(count, degree, seed), F.train_camera(72, 56), F.smooth_target(72, 56, 53)


@pytest.mark.parametrize("count,degree", [(300, 0), (1000, 3)])
def test_pose_gradient_matches_the_model_bit_for_bit(count, degree):
    cloud, cam, target = _case(count, degree)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, pose_gradients=True)) as t, \
         vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as plain:
        stats = t.compute_gradients(cam, target)
        g = t.gradients()
        off = plain.compute_gradients(cam, target)
        g_off = plain.gradients()
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_fixed(r)
    _, want = m.pose(r)
    got = np.array(stats.pose_gradient, np.float32)
    print(f"pose gradient {got}", flush=True)
    np.testing.assert_array_equal(_bits(got), _bits(want))
    assert np.any(got != 0.0)
    assert not any(off.pose_gradient)
    for name in GROUPS:
        np.testing.assert_array_equal(_bits(getattr(g, name)), _bits(getattr(g_off, name)))


def test_step_fills_the_same_pose_gradient_and_batches_leave_it_zero():
    cloud, cam, target = _case()
    options = vsplat.TrainerOptions(force_fixed_point=True, pose_gradients=True)
    with vsplat.Trainer(cloud, options) as a, vsplat.Trainer(cloud, options) as b:
        want = a.compute_gradients(cam, target).pose_gradient
        got = b.step(cam, target, vsplat.LearningRates()).pose_gradient
        batch = b.step_views([cam, cam], [target, target], vsplat.LearningRates())
    np.testing.assert_array_equal(_bits(np.array(got, np.float32)), _bits(np.array(want, np.float32)))
    assert all(not any(s.pose_gradient) for s in batch)


def test_pose_gradient_agrees_with_finite_differences():
    cloud, cam, target = _case()

    def loss_at(move):
        return O.loss(O.render(cloud, O.refined_pose(cam, move), O.RefOptions(no_cutoffs=True)), target, 0.2).loss

    numeric = []
    for k in range(6):
        up, down = np.zeros(6, np.float32), np.zeros(6, np.float32)
        up[k], down[k] = 1e-3, -1e-3
        numeric.append((loss_at(up) - loss_at(down)) / 2e-3)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(pose_gradients=True)) as t:
        got = np.array(t.compute_gradients(cam, target).pose_gradient, np.float64)
    b = np.array(numeric)
    cosine = float(got @ b / math.sqrt((got @ got) * (b @ b)))
    ratio = math.sqrt((got @ got) / (b @ b))
    print(f"pose cosine {cosine:.6f}, length ratio {ratio:.5f}", flush=True)
    assert cosine > 0.95 and abs(ratio - 1.0) < 0.3
