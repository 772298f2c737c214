"""Pins of the pose gradient's oracle and of refined_pose, on the CPU
(trainer.cpp:120-137, 630-637; train_gradient.comp:276-279, 317-337):

* refined_pose: the reference's own closed form (tests/test_train.cpp:744-751):
  a point turns and shifts, the eye maps to the origin; the library's C
  function equals the oracle's restatement;
* pose_totals: the group tree plus the ordered double sum equals the exact sum
  within float rounding;
* Oracle B's pose gradient equals the pose identity taken on Oracle A's FP64
  position gradients (dL/dv = sum Rv dL/dp, dL/dw = sum t x Rv dL/dp), and
  agrees with central differences of Oracle A's loss under refined_pose as
  the reference's test requires (tests/test_train.cpp:722-742: cosine above
  0.95, length within 30%; the identity leaves out how the camera's rotation
  turns each Gaussian's shape).
"""
import math

import numpy as np

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat


def _check_close(a, b, min_cosine, max_ratio_error):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    cosine = float(a @ b / math.sqrt((a @ a) * (b @ b)))
    ratio = math.sqrt((a @ a) / (b @ b))
    print(f"cosine {cosine:.6f}, length ratio {ratio:.5f}")
    assert cosine > min_cosine and abs(ratio - 1.0) < max_ratio_error


def test_refined_pose_turns_and_shifts_a_point():
    cam = F.train_camera(72, 56)
    for moved in (O.refined_pose(cam, [0.0, 0.1, 0.0, 0.2, 0.0, 0.0]),
                  vsplat.refined_pose(cam, [0.0, 0.1, 0.0, 0.2, 0.0, 0.0])):
        point = moved.view.astype(np.float64) @ np.array([0.0, 0.0, 3.0, 1.0])
        assert abs(point[0] - (3.0 * math.sin(0.1) + 0.2)) < 1e-5
        assert abs(point[2] - 3.0 * math.cos(0.1)) < 1e-5
        eye = moved.view.astype(np.float64) @ np.append(moved.position.astype(np.float64), 1.0)
        assert np.linalg.norm(eye[:3]) < 1e-5
    # No move is no change, to rounding.
    same = vsplat.refined_pose(cam, np.zeros(6, np.float32))
    np.testing.assert_allclose(same.view, cam.view, atol=1e-7)


def test_library_refined_pose_equals_the_oracle():
    rng = np.random.default_rng(3)
    cam = F.look_at([1.0, -0.5, -3.0], [0.1, 0.2, 0.0], [0.0, -1.0, 0.0], 50.0, 64, 48)
    for _ in range(20):
        move = rng.normal(0.0, 0.3, 6).astype(np.float32)
        if rng.random() < 0.3:
            move[:3] *= 1e-9  # the small-angle limits
        a, b = vsplat.refined_pose(cam, move), O.refined_pose(cam, move)
        np.testing.assert_allclose(a.view, b.view, rtol=0, atol=2e-6)
        np.testing.assert_allclose(a.position, b.position, rtol=0, atol=2e-6)
        # The rotation part stays a rotation.
        R = a.view[:3, :3].astype(np.float64)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)


def test_pose_totals_sum_the_terms():
    rng = np.random.default_rng(7)
    for n in (1, 255, 256, 257, 5000):
        terms = rng.normal(0.0, 1.0, (n, 6)).astype(np.float32)
        exact = np.array([math.fsum(terms[:, k].astype(np.float64)) for k in range(6)])
        got = O.pose_totals(terms).astype(np.float64)
        bound = 1e-6 * np.abs(terms).astype(np.float64).sum(0) + 1e-30
        assert np.all(np.abs(got - exact) <= bound), n


def _case():
    cam = F.train_camera(72, 56)
    return F.random_cloud(300, 0, 59), cam, F.smooth_target(72, 56, 53)


def test_model_pose_equals_the_identity_on_fp64_position_gradients():
    cloud, cam, target = _case()
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_float(r)
    terms, totals = m.pose(r)
    ref_loss = O.loss(O.render(cloud, cam), target, 0.2)
    gp = O.gradients(cloud, cam, ref_loss.adjoint).positions
    V = cam.view.astype(np.float64)
    t = (V[:3, :3] @ cloud.positions.astype(np.float64).T).T + V[:3, 3]
    v = (V[:3, :3] @ gp.T).T
    want = np.concatenate([np.cross(t, v).sum(0), v.sum(0)])
    print(f"model {totals}, identity on FP64 gradients {want}")
    np.testing.assert_allclose(totals, want, rtol=0, atol=1e-4 * np.abs(want).max())
    assert np.count_nonzero(np.any(terms != 0, 1)) > 50


def test_model_pose_agrees_with_finite_differences():
    cloud, cam, target = _case()

    def loss_at(move):
        return O.loss(O.render(cloud, O.refined_pose(cam, move), O.RefOptions(no_cutoffs=True)), target, 0.2).loss

    numeric = []
    for k in range(6):
        up, down = np.zeros(6, np.float32), np.zeros(6, np.float32)
        up[k], down[k] = 1e-3, -1e-3
        numeric.append((loss_at(up) - loss_at(down)) / 2e-3)
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_float(r)
    _, totals = m.pose(r)
    print(f"model {totals}, finite differences {np.array(numeric)}")
    _check_close(totals, numeric, 0.95, 0.3)
