"""Pins of the oracle's position noise and opacity cap (train_noise.comp,
trainer.cpp:806-819), on the CPU:

* the reference's own case (tests/test_train.cpp:322-350): identity
  covariance, transparent Gaussians move with mean 0 and standard deviation
  scaler / (1 + e^-0.5), opaque ones stay, a degenerate axis stays finite;
* the uniforms lie in (0, 1] with the moments of U(0, 1], and differ between
  iterations; the Box-Muller normals have mean 0 and identity covariance;
* a rotated, anisotropic Gaussian's moves have covariance Sigma^2 with
  Sigma = R diag(exp(2 s)) R^T, which a transposed rotation fails;
* the gate: logistic(100 ((1 - opacity) - 0.995)), nothing below 1e-6;
* the opacity cap: sigmoid(limit) is the maximum.
"""
import math

import numpy as np

from oracle import oracle as O


def _quat_matrix(q):
    w, x, y, z = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _cloud(count, opacity=-20.0, scales=(0.0, 0.0, 0.0), rotation=(1.0, 0.0, 0.0, 0.0)):
    return (np.zeros((count, 3), np.float32), np.tile(np.float32(scales), (count, 1)),
            np.tile(np.float32(rotation), (count, 1)), np.full(count, opacity, np.float32))


def test_reference_case_transparent_move_and_opaque_stay():
    n = 4000
    p, s, r, o = _cloud(n)
    o[n // 2:] = 20.0
    s[n - 1, 0] = -30.0  # a degenerate axis still gives a finite position
    moved = O.position_noise(p, s, r, o, 0.5, 7)
    x = moved[: n // 2, 0]
    print(f"noise on transparent Gaussians: mean {x.mean():.4f}, standard deviation {x.std():.4f}")
    assert abs(x.mean()) < 0.05
    assert abs(x.std() - 0.5 / (1.0 + math.exp(-0.5))) < 0.03
    assert np.all(moved[n // 2:] == 0.0)
    assert np.all(np.isfinite(moved))


def test_uniforms_lie_in_the_half_open_interval_with_uniform_moments():
    u = O.noise_uniforms(200_000, 3).astype(np.float64)
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 3e-3
    assert abs(u.var() - 1.0 / 12.0) < 2e-3
    # The four streams are uncorrelated, and another iteration draws anew.
    c = np.corrcoef(u.T)
    assert np.all(np.abs(c - np.eye(4)) < 0.01)
    v = O.noise_uniforms(200_000, 4).astype(np.float64)
    assert np.mean(u == v) < 1e-3
    assert abs(np.corrcoef(u[:, 0], v[:, 0])[0, 1]) < 0.01
    np.testing.assert_array_equal(O.noise_uniforms(1000, 3), O.noise_uniforms(1000, 3))


def test_pcg_hash_is_injective_on_a_sample():
    x = np.random.default_rng(5).choice(2**32, size=1 << 20, replace=False).astype(np.uint64)
    assert np.unique(O.pcg_hash(x)).size == x.size


def test_normals_have_zero_mean_and_identity_covariance():
    n = 200_000
    p, s, r, o = _cloud(n)
    gate = 1.0 / (1.0 + math.exp(-0.5))
    normals = O.position_noise(p, s, r, o, 1.0, 11) / gate
    assert np.all(np.abs(normals.mean(0)) < 0.01)
    assert np.all(np.abs(np.cov(normals.T) - np.eye(3)) < 0.015)
    # Standard-normal tails: about 4.55% beyond two deviations.
    assert abs(np.mean(np.abs(normals) > 2.0) - 0.0455) < 0.003


def test_rotated_anisotropic_moves_have_covariance_sigma_squared():
    n = 200_000
    q = (0.8, 0.3, -0.4, 0.33)
    scales = (math.log(0.5), 0.0, math.log(1.5))
    p, s, r, o = _cloud(n, scales=scales, rotation=q)
    gate = 1.0 / (1.0 + math.exp(-0.5))
    moves = O.position_noise(p, s, r, o, 1.0, 5) / gate
    R = _quat_matrix(q)
    sigma = R @ np.diag(np.exp(2.0 * np.array(scales))) @ R.T
    want = sigma @ sigma
    got = np.cov(moves.T)
    scale = np.abs(want).max()
    print(f"covariance error {np.abs(got - want).max() / scale:.4f}")
    assert np.abs(got - want).max() < 0.02 * scale
    wrong = R.T @ np.diag(np.exp(2.0 * np.array(scales))) @ R
    assert np.abs(got - wrong @ wrong).max() > 0.2 * scale


def test_gate_is_the_steep_logistic():
    n = 100_000
    for logit in (-20.0, -5.2, -4.6, 0.0, 20.0):
        p, s, r, o = _cloud(n, opacity=logit)
        x = O.position_noise(p, s, r, o, 1.0, 2)
        opacity = 1.0 / (1.0 + math.exp(-logit))
        gate = 1.0 / (1.0 + math.exp(-100.0 * ((1.0 - opacity) - 0.995)))
        if gate < 1e-6:
            assert np.all(x == 0.0), logit
        else:
            assert abs(x.std() - gate) < 0.01 * gate, (logit, x.std(), gate)


def test_opacity_limit_is_the_logit():
    for m in (0.01, 0.05, 0.5, 0.9):
        limit = O.opacity_limit(m)
        assert abs(1.0 / (1.0 + math.exp(-limit)) - float(np.float32(m))) < 1e-12
