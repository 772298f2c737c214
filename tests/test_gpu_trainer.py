"""The trainer surface on the GPU, beyond the per-stage parity of test_gpu_parity.py:

* the Adam step (train_gradient.comp step mode, train_sh_step.comp) bit for bit
  against Oracle B, through vsplat_step, through compute_gradients +
  apply_gradients, and through the device-resident queue;
* degenerate Gaussians leave the render bitwise unchanged (tests/test_train.cpp:544-586);
* the full 1M-Gaussian, 1280x720 workload through properties that do not need
  the CPU model: sortedness and stability of both sorts, ranges partitioning the
  entries, fixed-point results identical across runs;
* running out of reserved tile entries in the queue, and the memory budget
  (BudgetExceeded, trainer.hpp:46-49);
* the prefetched next target (trainer.cpp:539-564) gives the same step as a
  directly passed one.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

RATES = vsplat.LearningRates(position=1.6e-4, scale=0.005, rotation=0.001, opacity=0.05, sh=0.025, sh_rest_scale=0.05)
RATE_TUPLE = (RATES.position, RATES.scale, RATES.rotation, RATES.opacity, RATES.sh)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _case(name):
    if name == "ref60":
        return F.random_cloud(60, 1, 13), F.train_camera(72, 56), F.smooth_target(72, 56, 11)
    if name == "deg3":
        return F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)
    if name == "config1_small":
        w = F.config(1, count=20_000)
        return w.cloud, w.cameras[0], w.targets[0]
    raise KeyError(name)


def _model_step(cloud, cam, target, sh_storage=0):
    m = O.Model(cloud, cam, sh_storage=sh_storage)
    r = m.forward()
    m.loss(r, target, 0.2)
    m.backward_fixed(r)
    return m.step(r, RATE_TUPLE, rest_scale=RATES.sh_rest_scale, step=1)


def _assert_cloud_equals(c, want, coefficients, sh_storage=0):
    pos, sc, rot, op, words = want
    np.testing.assert_array_equal(_bits(c.positions.reshape(-1)), _bits(pos))
    np.testing.assert_array_equal(_bits(c.scales.reshape(-1)), _bits(sc))
    np.testing.assert_array_equal(_bits(c.rotations.reshape(-1)), _bits(rot))
    np.testing.assert_array_equal(_bits(c.opacities), _bits(op))
    sh = F.unpack_sh(words, coefficients, sh_storage).reshape(c.sh.shape)
    np.testing.assert_array_equal(_bits(c.sh), _bits(sh))


@pytest.mark.parametrize("name,sh_storage", [("ref60", 0), ("deg3", 0), ("config1_small", 0), ("ref60", 1), ("deg3", 1)])
def test_adam_step_bit_exact(name, sh_storage):
    cloud, cam, target = _case(name)
    want = _model_step(cloud, cam, target, sh_storage)
    coefficients = (cloud.sh_degree + 1) ** 2
    options = vsplat.TrainerOptions(force_fixed_point=True, sh_storage=sh_storage)
    with vsplat.Trainer(cloud, options) as t:
        t.step(cam, target, RATES)
        stepped = t.cloud()
    _assert_cloud_equals(stepped, want, coefficients, sh_storage)
    moved = np.abs(stepped.opacities - cloud.opacities)
    assert moved.max() > 0.04  # the first step moves by about the rate
    # The same step as compute_gradients then apply_gradients (trainer.hpp:104, 121).
    with vsplat.Trainer(cloud, options) as t:
        t.compute_gradients(cam, target)
        t.apply_gradients(RATES)
        _assert_cloud_equals(t.cloud(), want, coefficients, sh_storage)
    # And through the device-resident queue.
    with vsplat.Trainer(cloud, options) as t, vsplat.DeviceBuffer(target) as d:
        t.compute_gradients(cam, target)  # sizes the tile-entry reservation, changes no parameter
        t.enqueue(cam, d.ptr, vsplat.MODE_STEP, RATES)
        t.synchronize()
        _assert_cloud_equals(t.cloud(), want, coefficients, sh_storage)


def test_degenerate_gaussians_do_not_change_the_render():
    """test_train.cpp:544-586: a NaN position, an overflowing scale, a quaternion
    shorter than 1e-4, a NaN opacity and a NaN SH coefficient."""
    cam = F.train_camera(72, 56)
    healthy = F.random_cloud(40, 1, 41)
    c = F.Cloud.zeros(45, 1)
    for k in ("positions", "scales", "rotations", "opacities", "sh"):
        getattr(c, k)[:40] = getattr(healthy, k)
    for i in range(40, 45):
        c.positions[i, 2] = 4.0
        c.scales[i] = np.log(0.3)
        c.rotations[i, 0] = 1.0
        c.opacities[i] = 2.0
    c.positions[40, 0] = np.nan
    c.scales[41, 0] = 200.0
    c.rotations[42, 0] = 1e-5
    c.opacities[43] = np.nan
    c.sh[44, 0, 0] = np.nan
    with vsplat.Trainer(healthy) as a, vsplat.Trainer(c) as b:
        base, image = a.render(cam), b.render(cam)
    np.testing.assert_array_equal(_bits(image), _bits(base))


def test_full_size_workload_properties():
    """The 1M-Gaussian, 1280x720 workload: sorts, lists and ranges hold their
    invariants at full size, and the fixed-point pipeline gives the same bits twice."""
    w = F.config(1)
    cam, target = w.cameras[0], w.targets[0]
    with vsplat.Trainer(w.cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        stats = t.compute_gradients(cam, target)
        keys, values = t.inspect("keys"), t.inspect("values")
        ek, ev, ranges = t.inspect("entry_keys"), t.inspect("entry_values"), t.inspect("ranges")
        coverage = t.inspect("coverage")
        raw1 = t.inspect("raw").copy()
        g1 = t.gradients()
        image1 = t.inspect("image").copy()
        stats2 = t.compute_gradients(cam, target)
        raw2 = t.inspect("raw")
        g2 = t.gradients()
        image2 = t.inspect("image")
    n = w.cloud.count
    print(f"visible {stats.visible_splats} entries {stats.tile_entries} loss {stats.loss:.6f}")
    assert stats.visible_splats > 0.9 * n and stats.tile_entries > 5e6
    # Depth sort: a permutation, keys ascending, equal keys in index order.
    assert np.array_equal(np.sort(values), np.arange(n, dtype=np.uint32))
    k = keys.astype(np.int64)
    assert (np.diff(k) >= 0).all()
    same = np.diff(k) == 0
    assert (np.diff(values.astype(np.int64))[same] > 0).all()
    assert int((keys != 0xFFFFFFFF).sum()) == stats.visible_splats
    E = stats.tile_entries
    assert len(ek) == E and len(ev) == E
    # Tile sort: tiles ascending, depth order kept within a tile.
    assert (np.diff(ek.astype(np.int64)) >= 0).all()
    rank = np.empty(n, np.int64)
    rank[values] = np.arange(n)
    within = np.diff(ek.astype(np.int64)) == 0
    assert (np.diff(rank[ev])[within] > 0).all()
    # Every splat appears as often as its tile count says; the buffer holds the
    # counts' exclusive scan (the emit offsets) in depth order.
    offsets = coverage[:n].astype(np.int64)
    assert offsets[0] == 0 and (np.diff(offsets) >= 0).all()
    counts = np.bincount(ev, minlength=n)
    assert np.array_equal(counts[values], np.diff(np.append(offsets, E)))
    # Ranges partition [0, E).
    nonempty = ranges[:, 1] > ranges[:, 0]
    assert ranges[nonempty, 1].astype(np.int64).sum() - ranges[nonempty, 0].astype(np.int64).sum() == E
    for tile in np.flatnonzero(nonempty)[:: max(1, nonempty.sum() // 200)]:
        a, b = ranges[tile]
        assert (ek[a:b] == tile).all()
    # Determinism of the fixed-point path.
    assert stats2.tile_entries == E and stats2.loss == stats.loss
    np.testing.assert_array_equal(raw1, raw2)
    np.testing.assert_array_equal(_bits(image1), _bits(image2))
    for a, b in zip((g1.positions, g1.scales, g1.rotations, g1.opacities, g1.sh),
                    (g2.positions, g2.scales, g2.rotations, g2.opacities, g2.sh)):
        np.testing.assert_array_equal(_bits(a), _bits(b))
    assert np.isfinite(image1).all() and np.isfinite(g1.positions).all() and np.isfinite(g1.sh).all()


def test_queued_step_that_does_not_fit_changes_nothing():
    cloud, cam, target = _case("ref60")
    want = _model_step(cloud, cam, target)
    options = vsplat.TrainerOptions(force_fixed_point=True)
    with vsplat.Trainer(cloud, options) as t, vsplat.DeviceBuffer(target) as d:
        before = t.cloud()
        # No reservation yet: both queued steps are skipped, and Adam's count with them.
        t.enqueue(cam, d.ptr, vsplat.MODE_STEP, RATES)
        t.enqueue(cam, d.ptr, vsplat.MODE_STEP, RATES)
        with pytest.raises(RuntimeError, match="tile entries"):
            t.synchronize()
        after = t.cloud()
        for a, b in zip((before.positions, before.scales, before.rotations, before.opacities, before.sh),
                        (after.positions, after.scales, after.rotations, after.opacities, after.sh)):
            np.testing.assert_array_equal(_bits(a), _bits(b))
        t.reserve_entries(1 << 16)
        t.enqueue(cam, d.ptr, vsplat.MODE_STEP, RATES)
        t.synchronize()
        # This is Adam's first step: the same bits as a fresh trainer's step.
        _assert_cloud_equals(t.cloud(), want, (cloud.sh_degree + 1) ** 2)


def test_memory_budget_raises_budget_exceeded():
    cloud, cam, target = _case("config1_small")
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(memory_budget=1)) as t:
        with pytest.raises(vsplat.BudgetExceeded):
            t.compute_gradients(cam, target)


def test_prefetched_target_gives_the_same_step():
    cloud, cam, target = _case("ref60")
    other = F.smooth_target(72, 56, 12)
    options = vsplat.TrainerOptions(force_fixed_point=True)
    with vsplat.Trainer(cloud, options) as a, vsplat.Trainer(cloud, options) as b:
        a.step(cam, target, RATES, next_target=other)
        a.step(cam, other, RATES)
        b.step(cam, target, RATES)
        b.step(cam, other.copy(), RATES)
        ca, cb = a.cloud(), b.cloud()
    for x, y in zip((ca.positions, ca.opacities, ca.sh), (cb.positions, cb.opacities, cb.sh)):
        np.testing.assert_array_equal(_bits(x), _bits(y))


def test_save_ply_writes_the_trainers_cloud(tmp_path):
    """vsplat_save_ply: the trainer's cloud after a step (SH as stored) written
    as a 3DGS .ply reads back bit for bit."""
    cloud, cam, target = _case("deg3")
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        t.step(cam, target, RATES)
        want = t.cloud()
        t.save_ply(tmp_path / "trained.ply")
    got = vsplat.load_ply(tmp_path / "trained.ply")
    for name in ("positions", "scales", "rotations", "opacities", "sh"):
        np.testing.assert_array_equal(_bits(getattr(got, name)), _bits(getattr(want, name)))


@pytest.mark.parametrize("sh_storage", [0, 1])
def test_parameter_buffers_hold_the_cloud_in_the_render_layout(sh_storage):
    """Trainer::parameters (trainer.hpp:137; the viewer reads these buffers in
    place, viewer.cpp:326-337): device pointers and sizes of the five groups,
    positions to opacities as floats and the SH packed as half_sh_words
    (gaussian_cloud.hpp:40-49) or as bf16 words; after a step they hold what
    cloud() reads."""
    import ctypes as C
    cloud, cam, target = _case("deg3")
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, sh_storage=sh_storage)) as t:
        t.step(cam, target, RATES)
        c = t.cloud()
        groups = []
        for g in range(5):
            ptr, size = t.parameters(g)
            out = np.zeros(size // 4, np.uint32)
            assert vsplat.lib().vsplat_device_read(C.c_void_p(ptr), out.ctypes.data_as(C.c_void_p), out.nbytes) == 0
            groups.append(out)
        with pytest.raises(ValueError):
            t.parameters(5)
    n = cloud.count
    for got, want in zip(groups[:4], (c.positions, c.scales, c.rotations, c.opacities)):
        np.testing.assert_array_equal(got[: want.size], _bits(want).reshape(-1))
    words = F.pack_sh(c.sh, sh_storage).reshape(-1)
    assert groups[4].size >= words.size and words.size == n * F.sh_words((cloud.sh_degree + 1) ** 2, sh_storage)
    np.testing.assert_array_equal(groups[4][: words.size], words)
