"""The state gather of a changed cloud (train_gather.comp:16-24, trainer.cpp:749-787),
as the oracle restates it, pinned on the CPU: against a brute-force loop, the
identity table, the composition of two tables, fresh state for negative
sources and the rejection of out-of-range sources; and the C interface's
argument checks that need no GPU."""
import numpy as np
import pytest

from oracle import oracle as O


def _loop_gather(old, source, stride):
    out = []
    for s in source:
        for k in range(stride):
            out.append(0 if s < 0 else old[s * stride + k])
    return np.array(out, dtype=old.dtype)


@pytest.mark.parametrize("stride", [1, 3, 4, 48])
def test_gather_equals_a_loop(stride):
    rng = np.random.default_rng(stride)
    old = rng.integers(0, 2**32, size=17 * stride, dtype=np.uint64).astype(np.uint32)
    source = rng.integers(-3, 17, size=40)
    np.testing.assert_array_equal(O.gather(old, source, stride), _loop_gather(old, source, stride))


def test_identity_table_keeps_everything():
    old = (np.arange(30 * 4, dtype=np.uint64) * 2654435761 % 2**32).astype(np.uint32)
    np.testing.assert_array_equal(O.gather(old, np.arange(30), 4), old)


def test_two_tables_compose():
    rng = np.random.default_rng(5)
    old = rng.standard_normal(25 * 3).astype(np.float32)
    p = rng.integers(-1, 25, size=40)  # old -> middle
    q = rng.integers(-1, 40, size=33)  # middle -> new
    pq = np.where(q < 0, -1, p[np.maximum(q, 0)])  # a fresh middle Gaussian stays fresh
    np.testing.assert_array_equal(O.gather(O.gather(old, p, 3), q, 3), O.gather(old, pq, 3))


def test_negative_sources_start_fresh_and_sources_past_the_end_fail():
    old = np.ones(5 * 2, np.uint32)
    got = O.gather(old, [-1, 0, -7, 4], 2).reshape(4, 2)
    np.testing.assert_array_equal(got, [[0, 0], [1, 1], [0, 0], [1, 1]])
    with pytest.raises(ValueError):
        O.gather(old, [5], 2)
