"""Second-order training on the GPU (Trainer::enable_second_order,
Trainer::update_curvature, the trust-region step; trainer.cpp:642-667,
train_step.glsl:44-67, train_gradient.comp:131-143, train_hutchinson.comp):

* enable_second_order clears the curvature halves of every state but
  rotation's and keeps the first moments;
* update_curvature: the Hutchinson adjoint, the backward pass and the
  curvature chain equal Oracle B's bit for bit (fixed point, half and bf16 SH);
* a second-order step from that state equals Oracle B's bit for bit;
* the reference's own case (tests/test_train.cpp:753-785): opacity moves no
  more than its trust radius, and 300 steps with a curvature update every ten
  lower the loss.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

STATE = ("state_positions", "state_scales", "state_rotations", "state_opacities", "state_sh")
RATES = vsplat.LearningRates(position=1.6e-4, scale=0.005, rotation=0.001, opacity=0.05, sh=0.025, sh_rest_scale=0.05,
                             trust_region=1e-5)
RATE_TUPLE = (RATES.position, RATES.scale, RATES.rotation, RATES.opacity, RATES.sh)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1)


def _states(t):
    return [t.inspect(name).reshape(-1).copy() for name in STATE]


def _case():
    return F.random_cloud(120, 1, 67), F.train_camera(72, 56), F.random_target(72, 56, 61)


def test_enable_second_order_keeps_the_first_moments():
    cloud, cam, target = _case()
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        t.step(cam, target, vsplat.LearningRates())
        before = _states(t)
        t.enable_second_order()
        after = _states(t)
    for k, (b, a) in enumerate(zip(before, after)):
        assert np.any(b >> 16)
        np.testing.assert_array_equal(a, b if k == 2 else b & 0xFFFF)


@pytest.mark.parametrize("sh_storage", [0, 1])
def test_curvature_and_trust_region_step_match_the_model(sh_storage):
    cloud, cam, target = _case()
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, sh_storage=sh_storage)) as t:
        t.enable_second_order()
        t.compute_gradients(cam, target)
        stored = t.cloud()
        t.update_curvature()
        curved = _states(t)
        t.step(cam, target, RATES)
        after = t.cloud()
    cloud_s = stored
    # The curvature: Hutchinson adjoint through Oracle B's backward, then the curvature chain.
    m = O.Model(cloud_s, cam, sh_storage=sh_storage)
    r = m.forward()
    m.loss(r, target, 0.2)
    r.adjoint = O.hutchinson_adjoint(cam.width * cam.height, 1)
    # The curvature's sums carry no 1 / (3 W H): their own fixed-point scale.
    m.backward_fixed(r, scale=O.CURVATURE_FIXED_SCALE)
    n = cloud.count
    states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 3 * cloud_s.coefficients)]
    m.curvature(r, states, 1)
    print(f"curvature words set: {sum(int(np.count_nonzero(s >> 16)) for s in curved)}", flush=True)
    for got, want in zip(curved, states):
        np.testing.assert_array_equal(got, want)
    # The step from that state (Adam's step 1 for rotation, the trust region for the rest).
    r2 = m.forward()
    m.loss(r2, target, 0.2)
    m.backward_fixed(r2)
    pos, sc, rot, op, words = m.step(r2, RATE_TUPLE, rest_scale=RATES.sh_rest_scale, step=1, states=states, second_order=True,
                                     trust_region=RATES.trust_region, curvature_updates=1)
    np.testing.assert_array_equal(_bits(after.positions), _bits(pos))
    np.testing.assert_array_equal(_bits(after.scales), _bits(sc))
    np.testing.assert_array_equal(_bits(after.rotations), _bits(rot))
    np.testing.assert_array_equal(_bits(after.opacities), _bits(op))
    np.testing.assert_array_equal(_bits(after.sh), _bits(F.unpack_sh(words, cloud_s.coefficients, sh_storage)))


def test_second_order_steps_fit_a_view_within_their_trust_region():
    # tests/test_train.cpp:753-785
    cloud, cam, target = _case()
    with vsplat.Trainer(cloud, vsplat.TrainerOptions()) as t:
        t.enable_second_order()
        t.compute_gradients(cam, target)
        t.update_curvature()
        before = t.cloud()
        first = t.step(cam, target, RATES)
        after = t.cloud()
        a = np.maximum(1.0 / (1.0 + np.exp(-before.opacities.astype(np.float64))), 1e-3)
        radius = np.sqrt(4.0 * a * RATES.trust_region) / np.maximum(a * (1.0 - a), 1e-3)
        assert np.all(np.abs(after.opacities.astype(np.float64) - before.opacities) <= radius * 1.001)
        loss = first.loss
        for i in range(300):
            loss = t.step(cam, target, RATES).loss
            if i % 10 == 0:
                t.update_curvature()
        with pytest.raises(vsplat.Unsupported):
            t.apply_gradients(RATES)
    print(f"loss {first.loss:.6f} -> {loss:.6f}", flush=True)
    assert np.isfinite(loss) and loss < first.loss


def test_curvature_after_a_batch_takes_the_last_views_buffers():
    """After a batch of two sizes, update_curvature works on the last view
    (its camera, lists and image), bit for bit against Oracle B there; the
    larger last view must not be read through the first view's buffers."""
    cloud = F.random_cloud(120, 1, 67)
    cams = [F.train_camera(72, 56), F.train_camera(100, 75)]
    targets = [F.random_target(72, 56, 61), F.random_target(100, 75, 62)]
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        t.enable_second_order()
        t.step_views(cams, targets)  # gradients mode: no parameter changes
        t.update_curvature()
        curved = _states(t)
    m = O.Model(cloud, cams[1])
    r = m.forward()
    m.loss(r, targets[1], 0.2)
    r.adjoint = O.hutchinson_adjoint(cams[1].width * cams[1].height, 1)
    m.backward_fixed(r, scale=O.CURVATURE_FIXED_SCALE)
    n = cloud.count
    states = [np.zeros(k, np.uint32) for k in (n * 3, n * 3, n * 4, n, n * 3 * cloud.coefficients)]
    m.curvature(r, states, 1)
    assert sum(int(np.count_nonzero(s >> 16)) for s in curved) > 0
    for got, want in zip(curved, states):
        np.testing.assert_array_equal(got, want)


def test_curvature_needs_a_view_that_ran():
    """Nothing ran after an autotune: update_curvature refuses (a logic error)
    before any launch."""
    cloud, cam, target = _case()
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t, vsplat.DeviceBuffer(target) as d:
        t.enable_second_order()
        with pytest.raises(vsplat.LogicError):
            t.update_curvature()
        t.compute_gradients(cam, target)
        t.autotune_tiles(cam, d.ptr)
        with pytest.raises(vsplat.LogicError):
            t.update_curvature()
        t.compute_gradients(cam, target)
        t.update_curvature()
