"""Exposure and view scoring (Trainer::set_exposure, Trainer::score_view;
trainer.cpp:397-425, 519-532) in the oracles, on the CPU.

The float32 restatements of the elementwise passes (oracle.exposed_image,
exposure_backward, score_fields, error_mask, score_statistics) are pinned by
closed forms and by Oracle A: a neutral exposure is the identity; the exposure
gradient of the FP32 model equals FP64 central differences of Oracle A's loss
of the exposed image within the reference's tolerance (tests/test_train.cpp:
405-443); flat targets have no edge and a step edge has the closed-form
weight; the error mask's threshold; and the FP32 model's count of marked
pixels per Gaussian equals Oracle A's (reference.cpp:248-272) within the
reference's tolerance (tests/test_train.cpp:588-621), the edge score as the
reference's test expects it (tests/test_train.cpp:623-657)."""
import numpy as np

from oracle import oracle as O
from vsplat import fixtures as F

EXPOSURE = np.array([1.1, 0.05, -0.02, 0.03, 0.02, 0.9, 0.04, -0.01, -0.03, 0.06, 1.05, 0.02], np.float32)


def test_neutral_exposure_is_the_identity():
    rng = np.random.default_rng(1)
    image = rng.random((50, 4)).astype(np.float32)
    adjoint = rng.standard_normal((50, 4)).astype(np.float32)
    np.testing.assert_array_equal(O.exposed_image(image, O.NEUTRAL_EXPOSURE), image)
    back, totals = O.exposure_backward(image, adjoint, O.NEUTRAL_EXPOSURE)
    np.testing.assert_array_equal(back, adjoint)
    want = np.concatenate([adjoint[:, :3, None] * np.concatenate([image[:, :3], np.ones((50, 1), np.float32)], 1)[:, None, :]],
                          0).reshape(50, 12).astype(np.float64).sum(0)
    np.testing.assert_allclose(totals.astype(np.float64).sum(0), want, rtol=1e-5, atol=1e-7)


def _exposed_loss(image, transform, target, lam):
    t = np.asarray(transform, np.float64).reshape(3, 4)
    rgb = image.rgb @ t[:, :3].T + t[:, 3]
    return O.loss(O.RefImage(image.width, image.height, rgb, image.log_t, image.stop), target, lam).loss


def test_exposure_gradient_matches_fp64_differences():
    cloud, cam, target = F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.smooth_target(72, 56, 11)
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2, exposure=EXPOSURE)
    got = r.exposure_gradient.astype(np.float64)
    image = O.render(cloud, cam)
    numeric = np.zeros(12)
    for k in range(12):
        up, down = EXPOSURE.astype(np.float64).copy(), EXPOSURE.astype(np.float64).copy()
        up[k] += 1e-4
        down[k] -= 1e-4
        numeric[k] = (_exposed_loss(image, up, target, 0.2) - _exposed_loss(image, down, target, 0.2)) / 2e-4
    cosine = got @ numeric / (np.linalg.norm(got) * np.linalg.norm(numeric))
    assert cosine > 0.999, cosine
    assert abs(np.linalg.norm(got) / np.linalg.norm(numeric) - 1.0) < 0.02


def test_edge_weight_of_flat_and_stepped_targets():
    w, h = 24, 10
    flat = np.full((h, w, 4), 128, np.uint8)
    _, _, edge = O.score_fields(np.zeros((w * h, 4), np.float32), flat, w, h)
    assert not np.any(edge)
    step = np.full((h, w, 4), 255, np.uint8)
    step[:, : w // 2, :3] = 40
    step[:, w // 2:, :3] = 220
    _, _, edge = O.score_fields(np.zeros((w * h, 4), np.float32), step, w, h)
    edge = edge.reshape(h, w)
    lum = lambda v: (v / 255.0) * (0.299 + 0.587 + 0.114)
    want = abs(lum(220) - lum(40)) / 4.0
    np.testing.assert_allclose(edge[:, w // 2 - 1], want, rtol=1e-5)
    np.testing.assert_allclose(edge[:, w // 2], want, rtol=1e-5)
    assert not np.any(edge[:, : w // 2 - 1]) and not np.any(edge[:, w // 2 + 1:])


def test_error_and_its_mask():
    w, h = 4, 1
    image = np.zeros((4, 4), np.float32)
    image[:, 0] = [0.0, 0.25, 0.5, 1.0]
    target = np.zeros((h, w, 4), np.uint8)
    error, worst, _ = O.score_fields(image, target, w, h)
    np.testing.assert_allclose(error, [0.0, 0.25 / 3, 0.5 / 3, 1.0 / 3], rtol=1e-6)
    assert worst == int(np.float32(np.float32(1.0) / np.float32(3.0)) * np.float32(65535.0))
    mask = O.error_mask(error, worst, 0.5)
    np.testing.assert_array_equal(mask, [0, 0, 1, 1])
    assert not np.any(O.error_mask(error, 0, 0.5))


def _scored_model(cloud, cam, target, share=0.5, pixel_weight=2.0):
    m = O.Model(cloud, cam)
    r = m.forward()
    m.loss(r, target, 0.2, score_share=share)
    m.backward_float(r)
    stats = m.statistics(r)
    totals = r.tile_totals.astype(np.float64).sum(0) / (3.0 * cam.width * cam.height)
    loss = np.float32(0.8 * totals[1] + 0.2 * (1.0 - totals[0]))
    return O.score_statistics(stats, r.raw, 3.0 * cam.width * cam.height * pixel_weight, loss), loss


def test_marked_pixels_against_fp64():
    cloud, cam = F.random_cloud(60, 1, 47), F.train_camera(72, 56)
    target = F.random_target(72, 56, 43)
    cpu = O.marked_pixels(cloud, cam, target, 0.5)
    assert cpu.sum() > 0
    stats, loss = _scored_model(cloud, cam, target)
    # A pixel either way: the worst error is kept to 1/65535 (tests/test_train.cpp:611-615).
    np.testing.assert_allclose(stats[:, 5], 2.0 * cpu, atol=3.0)
    np.testing.assert_allclose(stats[:, 6], 2.0 * cpu * float(loss), atol=3.0)


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
    stats, _ = _scored_model(cloud, cam, target, pixel_weight=1.0)
    assert stats[:8, 7].sum() > 0.0
    assert stats[8:, 7].sum() == 0.0
    assert np.all(stats[:, 5] > 0.0)


def test_hutchinson_signs_known_answer():
    """train_hutchinson.comp's hash (a 32-bit integer mixer): known values
    worked by hand for seed 1, and signs roughly balanced."""
    def mix(x):
        x &= 0xFFFFFFFF
        x ^= x >> 16
        x = (x * 0x7FEB352D) & 0xFFFFFFFF
        x ^= x >> 15
        x = (x * 0x846CA68B) & 0xFFFFFFFF
        x ^= x >> 16
        return x
    adj = O.hutchinson_adjoint(4096, 1)
    for i in (0, 1, 777, 4095):
        bits = mix(i ^ mix(1))
        assert list(adj[i, :3]) == [1.0 if bits & (1 << c) else -1.0 for c in range(3)]
    assert abs(adj[:, :3].mean()) < 0.05 and not np.any(adj[:, 3])
