"""Trainer::apply_changes on the GPU (trainer.cpp:749-787), the device side of
densify, relocate and the remove_* calls:

* after two Adam steps, a change that prunes, clones and adds Gaussians keeps
  each survivor's optimiser state (the oracle's gather of train_gather.comp,
  bit for bit), gives new Gaussians zero moments, takes the values handed in,
  and clears the densification statistics; the next Adam step from the
  gathered state equals Oracle B's, bit for bit (half and bf16 SH storage);
* with no cloud handed in, the values are gathered on the device too: a pruned,
  permuted and partly duplicated cloud equals the old one through the table,
  and its fixed-point gradients equal those of a new trainer made from that
  cloud;
* bad tables fail with the reference's errors and leave the trainer as it
  was; a captured batch is dropped.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

RATES = vsplat.LearningRates(position=1.6e-4, scale=0.005, rotation=0.001, opacity=0.05, sh=0.025, sh_rest_scale=0.05)
RATE_TUPLE = (RATES.position, RATES.scale, RATES.rotation, RATES.opacity, RATES.sh)
STATE = ("state_positions", "state_scales", "state_rotations", "state_opacities", "state_sh")
GROUPS = ("positions", "scales", "rotations", "opacities", "sh")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1)


def _case():
    return F.random_cloud(300, 3, 7), F.train_camera(72, 56), F.smooth_target(72, 56, 17)


def _strides(cloud):
    return (3, 3, 4, 1, 3 * cloud.coefficients)


def _states(t):
    return [t.inspect(name).reshape(-1).copy() for name in STATE]


def _select(cloud, source):
    c = cloud.copy()
    for name in GROUPS:
        setattr(c, name, np.ascontiguousarray(getattr(cloud, name)[source]))
    return c


def _changes(old, seed):
    """Prune every seventh Gaussian, clone every eleventh (the clone moved a
    little, as densify_cloud's clones are), and add 25 new ones."""
    n = old.count
    keep = [i for i in range(n) if i % 7 != 3]
    clones = [i for i in range(n) if i % 11 == 0]
    fresh = F.random_cloud(25, old.sh_degree, seed)
    new = _select(old, np.array(keep + clones))
    new.positions[len(keep):] += np.float32(0.01)
    for name in GROUPS:
        setattr(new, name, np.ascontiguousarray(np.concatenate([getattr(new, name), getattr(fresh, name)])))
    return np.array(keep + clones + [-1] * 25, np.int64), new


@pytest.mark.parametrize("sh_storage", [0, 1])
def test_changes_keep_the_survivors_state_and_the_next_step_matches(sh_storage):
    cloud, cam, target = _case()
    options = vsplat.TrainerOptions(force_fixed_point=True, sh_storage=sh_storage)
    with vsplat.Trainer(cloud, options) as t:
        t.step(cam, target, RATES)
        t.step(cam, target, RATES)
        old_states = _states(t)
        assert all(np.any(s != 0) for s in old_states)
        source, new = _changes(t.cloud(), 99)
        t.apply_changes(new, source)
        assert t.count == len(source) == t.gaussian_count()

        states = _states(t)
        for got, old, stride in zip(states, old_states, _strides(cloud)):
            np.testing.assert_array_equal(got, O.gather(old, source, stride))
        stored = t.cloud()
        for name in GROUPS[:4]:
            np.testing.assert_array_equal(_bits(getattr(stored, name)), _bits(getattr(new, name)))
        words = F.pack_sh(new.sh, sh_storage)
        np.testing.assert_array_equal(_bits(stored.sh), _bits(F.unpack_sh(words, cloud.coefficients, sh_storage)))
        assert not np.any(t.densify_stats())

        # Adam's third step, from the gathered state.
        t.compute_gradients(cam, target)
        g = t.gradients()
        grads = [np.ascontiguousarray(getattr(g, name), np.float32).reshape(-1) for name in GROUPS]
        want = O.adam_from_gradients(stored, words, grads, RATE_TUPLE, rest_scale=RATES.sh_rest_scale, step=3,
                                     sh_storage=sh_storage, states=states)
        t.apply_gradients(RATES)
        after = t.cloud()
    pos, sc, rot, op, w = want
    np.testing.assert_array_equal(_bits(after.positions), _bits(pos))
    np.testing.assert_array_equal(_bits(after.scales), _bits(sc))
    np.testing.assert_array_equal(_bits(after.rotations), _bits(rot))
    np.testing.assert_array_equal(_bits(after.opacities), _bits(op))
    np.testing.assert_array_equal(_bits(after.sh), _bits(F.unpack_sh(w, cloud.coefficients, sh_storage)))


def test_device_gather_of_values_prunes_and_permutes():
    cloud, cam, target = _case()
    options = vsplat.TrainerOptions(force_fixed_point=True)
    rng = np.random.default_rng(3)
    with vsplat.Trainer(cloud, options) as t:
        t.step(cam, target, RATES)
        old, old_states = t.cloud(), _states(t)
        perm = rng.permutation(cloud.count)
        source = np.concatenate([perm[: cloud.count - 40], perm[:10]])  # 40 pruned, 10 copied twice
        t.apply_changes(None, source)
        got = t.cloud()
        for name in GROUPS:
            np.testing.assert_array_equal(_bits(getattr(got, name)), _bits(getattr(old, name)[source]))
        for s, o, stride in zip(_states(t), old_states, _strides(cloud)):
            np.testing.assert_array_equal(s, O.gather(o, source, stride))
        t.compute_gradients(cam, target)
        g = t.gradients()
        stats = t.densify_stats()
        with vsplat.Trainer(got, options) as fresh:
            fresh.compute_gradients(cam, target)
            want, want_stats = fresh.gradients(), fresh.densify_stats()
    for name in GROUPS:
        np.testing.assert_array_equal(_bits(getattr(g, name)), _bits(getattr(want, name)))
    np.testing.assert_array_equal(_bits(stats), _bits(want_stats))


def test_bad_tables_fail_and_change_nothing():
    cloud, cam, target = _case()
    with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
        t.step(cam, target, RATES)
        before, states = t.cloud(), _states(t)
        with pytest.raises(ValueError):
            t.apply_changes(None, [0, 1, cloud.count])  # past the old cloud
        with pytest.raises(ValueError):
            t.apply_changes(None, [0, -1])  # a new Gaussian needs values
        with pytest.raises(ValueError):
            t.apply_changes(_select(before, np.arange(3)), np.arange(4))  # counts differ
        with pytest.raises(RuntimeError):
            t.apply_changes(None, np.zeros(0, np.int64))  # "densification removed every Gaussian"
        assert t.count == cloud.count
        after = t.cloud()
        for name in GROUPS:
            np.testing.assert_array_equal(_bits(getattr(after, name)), _bits(getattr(before, name)))
        for s, o in zip(_states(t), states):
            np.testing.assert_array_equal(s, o)


def test_a_captured_batch_is_dropped():
    cloud, cam, target = _case()
    buffer = vsplat.DeviceBuffer(target)
    try:
        with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True)) as t:
            t.step_views([cam], [target], RATES)
            t.capture_views([cam], [buffer.ptr], vsplat.MODE_STEP, RATES)
            t.apply_changes(None, np.arange(cloud.count - 5))
            with pytest.raises(vsplat.LogicError):
                t.launch_captured()
            stats = t.step(cam, target, RATES)  # the trainer works on at its new size
            assert stats.visible_splats > 0
    finally:
        buffer.free()
