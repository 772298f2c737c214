"""Position noise and the opacity cap on the GPU (Trainer::add_position_noise,
Trainer::reset_opacity, TrainerOptions::positionNoise; trainer.cpp:601,
806-819, train_noise.comp):

* add_position_noise equals the oracle's float64 restatement within 1e-4 of
  each Gaussian's move (the kernel's cos, sin, exp and log are float32 library calls)
  and the reference's own case (tests/test_train.cpp:322-350);
* a step with position_noise > 0 (one view, and a batch of views) leaves
  every other parameter as a step without noise does, bit for bit, and moves
  the positions by the noise of the step just taken (iteration 1, scaler
  position_noise * rates.position) on the stepped cloud;
* replays of a captured batch take the noise of each step in turn (the
  iteration is the step count on the device), against the oracle;
* reset_opacity caps the opacities at the logit of the maximum, leaves lower
  ones bit for bit, clears the opacity's optimiser state and nothing else
  (tests/test_densify.cpp:619-622).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

OTHERS = ("scales", "rotations", "opacities", "sh")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1)


def _spread_cloud(count, seed):
    """A random cloud whose opacity logits span opaque to fully transparent."""
    cloud = F.random_cloud(count, 1, seed)
    cloud.opacities[:] = np.linspace(-12.0, 4.0, count, dtype=np.float32)[np.random.default_rng(seed).permutation(count)]
    return cloud


def _assert_noise(got, before, cloud_after, scaler, iteration):
    want = O.position_noise(before, cloud_after.scales, cloud_after.rotations, cloud_after.opacities, scaler, iteration)
    move = np.abs(want - np.asarray(before, np.float64).reshape(-1, 3))
    # Relative to the whole move: a normal near zero carries the float32
    # rounding of 2 pi u, small against the move but not against itself.
    tol = 1e-4 * move.max(1, keepdims=True) + 4e-7 * np.maximum(np.abs(want), 1.0)
    err = np.abs(np.asarray(got, np.float64).reshape(-1, 3) - want)
    print(f"moved {np.mean(move.max(1) > 0):.3f} of the Gaussians, worst error / tolerance {np.max(err / tol):.3f}", flush=True)
    assert np.all(err <= tol)
    return move


def test_add_position_noise_matches_the_oracle():
    cloud = _spread_cloud(3000, 21)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions()) as t:
        stored = t.cloud()  # SH as the trainer stores it (half floats)
        t.add_position_noise(0.3, 17)
        got = t.cloud()
    for name in OTHERS:
        np.testing.assert_array_equal(_bits(getattr(got, name)), _bits(getattr(stored, name)))
    move = _assert_noise(got.positions, cloud.positions, cloud, 0.3, 17)
    assert np.mean(move.max(1) > 0) > 0.3
    opaque = 1.0 / (1.0 + np.exp(-cloud.opacities.astype(np.float64))) > 0.5
    np.testing.assert_array_equal(_bits(got.positions[opaque]), _bits(cloud.positions[opaque]))


def test_reference_case_transparent_move_and_opaque_stay():
    n = 4000
    cloud = F.Cloud.zeros(n, 0)
    cloud.rotations[:, 0] = 1.0
    cloud.opacities[: n // 2] = -20.0
    cloud.opacities[n // 2:] = 20.0
    cloud.scales[n - 1, 0] = -30.0
    with vsplat.Trainer(cloud, vsplat.TrainerOptions()) as t:
        t.add_position_noise(0.5, 7)
        moved = t.cloud().positions
    x = moved[: n // 2, 0].astype(np.float64)
    print(f"noise on transparent Gaussians: mean {x.mean():.4f}, standard deviation {x.std():.4f}", flush=True)
    assert abs(x.mean()) < 0.05
    assert abs(x.std() - 0.5 / (1.0 + math.exp(-0.5))) < 0.03
    assert np.all(moved[n // 2:] == 0.0)
    assert np.all(np.isfinite(moved))


@pytest.mark.parametrize("batched", [False, True])
def test_step_with_noise_moves_the_stepped_positions(batched):
    cloud = _spread_cloud(400, 5)
    cam, target = F.train_camera(72, 56), F.smooth_target(72, 56, 3)
    rates = vsplat.LearningRates()
    noise = 5e4
    results = []
    for option in (noise, 0.0):
        # Fixed-point sums, so that the two runs' gradients agree bit for bit.
        with vsplat.Trainer(cloud, vsplat.TrainerOptions(position_noise=option, force_fixed_point=True)) as t:
            if batched:
                t.step_views([cam], [target], rates)
            else:
                t.step(cam, target, rates)
            results.append(t.cloud())
    noisy, plain = results
    for name in OTHERS:
        np.testing.assert_array_equal(_bits(getattr(noisy, name)), _bits(getattr(plain, name)))
    move = _assert_noise(noisy.positions, plain.positions, plain, noise * rates.position, 1)
    assert np.mean(move.max(1) > 0) > 0.3


def test_reset_opacity_caps_and_clears_the_opacity_state():
    cloud = _spread_cloud(500, 9)
    cam, target = F.train_camera(72, 56), F.smooth_target(72, 56, 3)
    with vsplat.Trainer(cloud, vsplat.TrainerOptions()) as t:
        t.step(cam, target, vsplat.LearningRates())
        before = t.cloud()
        states = {g: t.inspect(f"state_{g}").copy() for g in ("positions", "scales", "rotations", "opacities")}
        assert np.any(states["opacities"] != 0)
        t.reset_opacity(0.01)
        after = t.cloud()
        cleared = t.inspect("state_opacities")
        kept = {g: t.inspect(f"state_{g}") for g in ("positions", "scales", "rotations")}
        with pytest.raises(ValueError):
            t.reset_opacity(1.0)
    limit = O.opacity_limit(0.01)
    capped = after.opacities > np.float32(limit) - np.float32(1e-5)
    high = before.opacities.astype(np.float64) > limit
    assert np.any(high) and np.any(~high)
    np.testing.assert_array_equal(_bits(after.opacities[~capped]), _bits(before.opacities[~capped]))
    values = np.unique(after.opacities[high])
    assert values.size == 1
    assert abs(float(values[0]) - limit) <= 2.0 * abs(float(np.spacing(np.float32(limit))))
    assert np.all(1.0 / (1.0 + np.exp(-after.opacities.astype(np.float64))) <= 0.0101)
    assert not np.any(cleared)
    for g, v in kept.items():
        np.testing.assert_array_equal(v, states[g])
    for name in ("positions", "scales", "rotations", "sh"):
        np.testing.assert_array_equal(_bits(getattr(after, name)), _bits(getattr(before, name)))


def test_captured_steps_take_the_noise_of_each_step():
    """Replays of a captured batch read the step count on the device: with every
    rate zero but a vanishing position rate (its Adam update rounds away), the
    positions after three replays are the noise of iterations 1, 2 and 3 added
    in turn, and nothing else changes."""
    cloud = _spread_cloud(400, 11)
    cam, target = F.train_camera(72, 56), F.smooth_target(72, 56, 3)
    rates = vsplat.LearningRates(position=1e-30, scale=0.0, rotation=0.0, opacity=0.0, sh=0.0, sh_rest_scale=0.05)
    options = vsplat.TrainerOptions(force_fixed_point=True, position_noise=1e30)
    with vsplat.DeviceBuffer(target) as d, vsplat.Trainer(cloud, options) as t:
        stored = t.cloud()
        t.reserve_entries(1 << 16)
        t.capture_views([cam], [d.ptr], vsplat.MODE_STEP, rates)
        t.launch_captured()
        t.synchronize_views()
        t.launch_captured(2)
        t.synchronize_views()
        got = t.cloud()
        t.release_captured()
    for name in OTHERS:
        np.testing.assert_array_equal(_bits(getattr(got, name)), _bits(getattr(stored, name)))
    scaler = float(np.float32(1e30) * np.float32(1e-30))
    want = stored.positions.astype(np.float32)
    for iteration in (1, 2, 3):
        want = O.position_noise(want, stored.scales, stored.rotations, stored.opacities, scaler, iteration).astype(np.float32)
    move = np.abs(want.astype(np.float64) - stored.positions.astype(np.float64))
    tol = 1e-4 * move.max(1, keepdims=True) + 1e-6 * np.maximum(np.abs(want), 1.0)
    err = np.abs(got.positions.astype(np.float64) - want.astype(np.float64))
    print(f"moved {np.mean(move.max(1) > 0):.3f}, worst error / tolerance {np.max(err / tol):.3f}", flush=True)
    assert np.mean(move.max(1) > 0) > 0.3
    assert np.all(err <= tol)
