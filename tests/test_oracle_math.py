"""The single-precision exp and log that Oracle B (and, bit for bit, the GPU)
uses: pinned against double-precision libm by an ulp sweep, and on the
special values IEEE fixes."""
import numpy as np

from oracle import oracle as O


def _ulp_error(got, x, f):
    want = f(x.astype(np.float64))
    spacing = np.spacing(np.abs(want.astype(np.float32))).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / spacing


def test_exp_within_about_one_ulp():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-87, 88.7, 1_000_000), rng.uniform(-10, 0, 500_000)]).astype(np.float32)
    assert _ulp_error(O.exp32(x), x, np.exp).max() < 1.1


def test_exp_gradual_underflow_rounds_once():
    x = np.random.default_rng(1).uniform(-103.9, -87.4, 200_000).astype(np.float32)
    err = np.abs(O.exp32(x).astype(np.float64) - np.exp(x.astype(np.float64)))
    assert (err <= 2.0 ** -149).all()  # within one denormal step


def test_log_within_one_ulp():
    rng = np.random.default_rng(2)
    x = np.concatenate([np.exp(rng.uniform(-87, 88, 1_000_000)), rng.uniform(0.004, 0.996, 500_000)]).astype(np.float32)
    assert _ulp_error(O.log32(x), x, np.log).max() < 1.0
    sub = (np.arange(1, 2 ** 23, 97) * 2.0 ** -149).astype(np.float32)
    assert _ulp_error(O.log32(sub), sub, np.log).max() < 1.0


def test_special_values():
    with np.errstate(all="ignore"):
        x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 88.73, 89.0, 1e30, -104.5, -1e30], np.float32)
        e = O.exp32(x)
        assert e[0] == 1 and e[1] == 1 and e[2] == np.inf and e[3] == 0 and np.isnan(e[4])
        assert e[5] == np.inf and e[6] == np.inf and e[7] == np.inf and e[8] == 0 and e[9] == 0
        assert O.exp32(np.float32([0.0]))[0] == 1.0 and O.log32(np.float32([1.0]))[0] == 0.0
        lx = np.array([0.0, -0.0, np.inf, -1.0, np.nan, 1.0, np.e], np.float32)
        lg = O.log32(lx)
        assert lg[0] == -np.inf and lg[1] == -np.inf and lg[2] == np.inf and np.isnan(lg[3]) and np.isnan(lg[4])
        assert lg[5] == 0.0 and abs(lg[6] - 1.0) < 1e-7


def test_exp_of_log_round_trips():
    x = np.random.default_rng(3).uniform(0.01, 100.0, 100_000).astype(np.float32)
    back = O.exp32(O.log32(x))
    assert np.max(np.abs(back / x - 1)) < 4e-7
