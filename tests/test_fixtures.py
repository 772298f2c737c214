"""The seeded inputs: the reference's fixtures reproduced draw for draw
(std::mt19937 + libstdc++'s generate_canonical), the explicit-RNG workloads,
cameras, and the trainer's SH packing."""
import math

import numpy as np

from vsplat import fixtures as F


def test_mt19937_known_answer():
    # The C++ standard fixes the 10000th output of a default-seeded mt19937 (seed 5489).
    assert F.StdMt19937(5489).raw(10000)[-1] == 4123659995
    assert F.StdMt19937(5489).raw(1)[0] == 3499211612


def test_generate_canonical_float_is_one_draw_over_2_32():
    raw = F.StdMt19937(7).raw(1000)
    c = F.StdMt19937(7).canonical_float(1000)
    assert np.array_equal(c, np.minimum(raw.astype(np.float32) / np.float32(2 ** 32), np.nextafter(np.float32(1), np.float32(0))))
    assert (c >= 0).all() and (c < 1).all()


def test_random_cloud_follows_reference_helpers():
    c = F.random_cloud(500, 3, 6)
    assert c.count == 500 and c.sh.shape == (500, 16, 3)
    assert (np.abs(c.positions[:, 0]) <= 1.4).all() and (np.abs(c.positions[:, 1]) <= 1.1).all()
    assert (c.positions[:, 2] >= 3).all() and (c.positions[:, 2] <= 6).all()
    s = np.exp(c.scales.astype(np.float64))
    assert (s >= 0.06 - 1e-6).all() and (s <= 0.26 + 1e-6).all()
    assert (np.abs(c.opacities) <= 1.5).all() and (np.abs(c.sh[:, 1:]) <= 0.2).all()
    again = F.random_cloud(500, 3, 6)
    assert np.array_equal(c.positions, again.positions) and np.array_equal(c.sh, again.sh)


def test_explicit_rng_is_a_pure_function_of_seed_stream_index():
    a = F.draw(5, 2, 1000)
    assert np.array_equal(a[:10], F.draw(5, 2, 10))
    assert not np.array_equal(a[:10], F.draw(5, 3, 10)) and not np.array_equal(a[:10], F.draw(6, 2, 10))
    # SplitMix64's published first output from state 0 stepping once: 0xE220A8397B1DCDAF.
    assert int(F.draw(0, 0, 1)[0]) == 0xE220A8397B1DCDAF
    n = F.normal(11, 0, 200_000)
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1) < 0.01
    u = F.uniform(11, 1, 200_000)
    assert (u >= 0).all() and (u < 1).all() and abs(u.mean() - 0.5) < 0.005


def test_look_at_matches_camera_cpp():
    cam = F.look_at((0, 0, -3), (0, 0, 0), (0, -1, 0), 60.0, 1280, 720)
    R = cam.view[:3, :3].astype(np.float64)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6) and abs(np.linalg.det(R) - 1) < 1e-6
    eye = cam.view.astype(np.float64) @ np.array([0, 0, -3, 1.0])
    assert np.allclose(eye[:3], 0, atol=1e-6)
    target = cam.view.astype(np.float64) @ np.array([0, 0, 0, 1.0])
    assert np.allclose(target[:2], 0, atol=1e-6) and abs(target[2] - 3) < 1e-6
    assert abs(cam.fy - 360 / math.tan(math.radians(30))) < 1e-3 and cam.fx == cam.fy
    assert cam.cx == 640 and cam.cy == 360


def test_half_sh_packing_round_trips():
    c = F.random_cloud(64, 3, 2)
    words = F.pack_half_sh(c.sh)
    assert words.shape == (64, F.half_sh_words(16)) and F.half_sh_words(16) == 26 and F.half_sh_words(1) == 3
    back = F.unpack_half_sh(words, 16)
    assert np.array_equal(back[:, 0], c.sh[:, 0])  # DC stays fp32
    assert np.array_equal(back[:, 1:], c.sh[:, 1:].astype(np.float16).astype(np.float32))
    assert (words[:, -1] >> 16 == 0).all()  # 45 halves: the last word's upper half is empty


def test_configs_are_the_stated_workloads():
    w0 = F.config(0)
    assert w0.cloud.count == 10_000 and w0.cloud.sh_degree == 0 and w0.cameras[0].width == 256 and w0.lambda_ssim == 0
    q = w0.cloud.rotations.astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1, atol=1e-6)
    w1 = F.config(1, count=1000)
    assert w1.cloud.sh_degree == 3 and (w1.cameras[0].width, w1.cameras[0].height) == (1280, 720) and w1.fixed_point
    w3 = F.config(3, count=100)
    assert [(c.width, c.height) for c in w3.cameras] == F.RING_SIZES


def test_bf16_packing_rounds_to_nearest_even():
    # Known answers on the float32 bits: exact, ties to even both ways, carry into the exponent.
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.0, float.fromhex("0x1.FFp0")], np.float32)
    assert F.float_to_bf16_bits(x).tolist() == [0x3F80, 0x3F80, 0x3F82, 0x3F81, 0xC000, 0x4000]
    # Against the nearest of the two bracketing bf16 values, in float64.
    rng = np.random.default_rng(5)
    v = (rng.standard_normal(100_000) * 10.0 ** rng.uniform(-6, 6, 100_000)).astype(np.float32)
    bits = F.float_to_bf16_bits(v)
    got = (bits << np.uint32(16)).view(np.float32).astype(np.float64)
    down = (v.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32).astype(np.float64)
    up = ((v.view(np.uint32) & np.uint32(0xFFFF0000)) + np.uint32(0x10000)).view(np.float32).astype(np.float64)
    exact = v.astype(np.float64)
    assert np.all(np.abs(got - exact) <= np.minimum(np.abs(down - exact), np.abs(up - exact)))
    # Packing: two per word, low half first, and the round trip.
    c = F.random_cloud(64, 3, 9)
    words = F.pack_bf16_sh(c.sh)
    assert words.shape == (64, F.bf16_sh_words(16)) and F.bf16_sh_words(16) == 24 and F.bf16_sh_words(1) == 2
    flat = c.sh.reshape(64, -1)
    assert (words[:, 0] & 0xFFFF).tolist() == F.float_to_bf16_bits(flat[:, 0]).tolist()
    assert (words[:, 0] >> 16).tolist() == F.float_to_bf16_bits(flat[:, 1]).tolist()
    back = F.unpack_bf16_sh(words, 16)
    np.testing.assert_array_equal(F.pack_bf16_sh(back), words)
    assert np.abs(back - c.sh).max() <= np.abs(c.sh).max() * 2 ** -8
    # Degree 0: three coefficients in two words, the unused half zero.
    w0 = F.pack_bf16_sh(F.random_cloud(5, 0, 3).sh)
    assert w0.shape == (5, 2) and (w0[:, 1] >> 16 == 0).all()
