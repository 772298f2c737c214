"""Pins Oracle B's hash and stochastic roundings by what the mathematics fixes.

The reference holds no vectors for them, so the exact neighbour a given
(value, index, seed) rounds to is fixed only by the formula. What the training
depends on is pinned here, each with a check that a plausible mistake fails:

* bf16 stochastic rounding of Adam's second moment and of the curvature
  (bf16.glsl:16-23): a value rounds to one of its two bf16 neighbours, up in
  magnitude with probability equal to its distance from the lower one in
  units of the gap, so that it is exact on average. Over 2^18 draws the rate
  of each of 64 values lies within a binomial bound of that distance; an
  always-truncating or a round-to-nearest copy fails the same statistic.
* fp16 stochastic rounding of the stored SH after a step (train_step.glsl:
  70-80): the same for halves, normal and subnormal, either sign.
* the hash (train_types.glsl:156-160; PCG's RXS M XS output permutation
  after one LCG step, O'Neill 2014, as the GPU hash of Jarzynski and Olano
  2020): a bijection of 32 bits for a fixed seed (4M indices spread over the
  32-bit range and 4M consecutive ones, no collision, where a random function
  would give about 2000; an even multiplier in the LCG step collides), its
  bits balanced, its top 10 bits uniform by chi-square, consecutive indices
  and consecutive seeds uncorrelated.
* SH dropout's draw (train_types.glsl:165-170): the share of Gaussians dropped
  is the probability, within a binomial bound.
"""
import numpy as np
import pytest

from oracle import oracle as O

K = 4096  # draws per value


def _binomial_ok(hits, p, n, sigmas=6.0):
    return abs(hits / n - p) <= sigmas * np.sqrt(max(p * (1 - p), 0.0) / n) + 2.0 / n


def _bf16_values():
    """64 floats with known offsets into their bf16 gap: (x, lower bits, fraction)."""
    rng = np.random.default_rng(3)
    fractions = [1 / 65536, 1 / 256, 0.01, 0.1, 0.25, 0.3, 0.5, 0.6, 0.75, 0.9, 0.99, 65535 / 65536]
    out = []
    while len(out) < 64:
        high = int(rng.integers(0x0080, 0x7F00))  # normal bf16 magnitudes
        sign = 0x8000 if rng.random() < 0.5 else 0
        f = fractions[len(out) % len(fractions)]
        low = int(round(f * 65536))
        bits = ((high | sign) << 16) | low
        out.append((np.array([bits], np.uint32).view(np.float32)[0], high | sign, low / 65536))
    return out


def _up_rate_bf16(round_fn):
    """For each value, how often round_fn(x, index, seed) is the upper-magnitude neighbour."""
    rates = []
    for k, (x, lower, f) in enumerate(_bf16_values()):
        got = round_fn(np.full(K, x, np.float32), np.arange(K, dtype=np.uint32), 1000 + k)
        assert np.all((got == lower) | (got == lower + 1)), "not a neighbour"
        rates.append((int(np.count_nonzero(got == lower + 1)), f))
    return rates


def _unbiased(rates):
    return all(_binomial_ok(hits, f, K) for hits, f in rates)


def test_bf16_stochastic_rounding_is_unbiased():
    rates = _up_rate_bf16(O.bf16_stochastic)
    assert _unbiased(rates), [(h / K, f) for h, f in rates if not _binomial_ok(h, f, K)]
    # Exact on average: the mean of the rounded values against the values, in gap units.
    mean_error = np.mean([h / K - f for h, f in rates])
    assert abs(mean_error) < 6 / np.sqrt(K * len(rates))


def test_the_bf16_statistic_fails_plausible_mistakes():
    def bits(x):
        return np.ascontiguousarray(x, np.float32).view(np.uint32)

    def truncate(x, index, seed):
        return bits(x) >> 16

    def nearest(x, index, seed):
        b = bits(x)
        return (b + 0x7FFF + ((b >> 16) & 1)) >> 16

    def wrong_half(x, index, seed):  # the hash's high bits where its low 16 belong
        return (bits(x) + (O.train_hash(index, seed) >> 16)) >> 16

    for mistake in (truncate, nearest):
        assert not _unbiased(_up_rate_bf16(mistake)), mistake.__name__
    # The high half of a good hash is as uniform as the low: this one passes,
    # which is why the hash is pinned by its own statistics below.
    assert _unbiased(_up_rate_bf16(wrong_half))


def _half_values():
    """Values between two adjacent halves: (x, lower-magnitude half bits, fraction of the gap)."""
    rng = np.random.default_rng(5)
    fractions = [0.001, 0.1, 0.25, 0.5, 0.75, 0.9, 0.999]
    out = []
    lowers = [int(h) for h in rng.integers(0x0001, 0x7BFF, 40)] + [0x0000, 0x0001, 0x0200, 0x03FF, 0x0400, 0x3C00, 0x7BFE]
    for k, lower in enumerate(lowers):
        for sign in (0x0000, 0x8000):
            f = fractions[(k + sign // 0x8000) % len(fractions)]
            lo = float(np.array([lower], np.uint16).view(np.float16)[0])
            hi = float(np.array([lower + 1], np.uint16).view(np.float16)[0])
            x = np.float32(lo + f * (hi - lo))
            frac = (float(x) - lo) / (hi - lo)  # what the float32 value actually holds
            if not 0.0 < frac < 1.0:
                continue
            out.append((np.float32(-x) if sign else x, lower | sign, frac))
    return out


def test_half_stochastic_rounding_is_unbiased():
    checked = 0
    for k, (x, lower, f) in enumerate(_half_values()):
        got = O.half_stochastic(np.full(K, x, np.float32), np.arange(K, dtype=np.uint32), 77 + k)
        assert np.all((got == lower) | (got == lower + 1)), (x, lower)
        hits = int(np.count_nonzero(got == lower + 1))
        assert _binomial_ok(hits, f, K), (x, hits / K, f)
        checked += 1
    assert checked >= 80
    # Values a half holds exactly stay as they are.
    exact = np.array([0.0, -0.0, 1.0, -2.5, 65504.0, 6.103515625e-05, 5.960464477539063e-08], np.float32)
    for k, v in enumerate(exact):
        got = O.half_stochastic(np.full(64, v, np.float32), np.arange(64, dtype=np.uint32), k)
        assert np.all(got == np.array([v], np.float16).view(np.uint16)[0])


def _spread_indices():
    """About 4M distinct indices drawn over the whole 32-bit range."""
    rng = np.random.default_rng(11)
    return np.unique(rng.integers(0, 1 << 32, 1 << 22, dtype=np.uint64)).astype(np.uint32)


@pytest.mark.parametrize("seed", [0, 12345, 0xFFFFFFFF])
def test_hash_is_a_bijection_with_balanced_uniform_bits(seed):
    spread = _spread_indices()
    assert np.unique(O.train_hash(spread, seed)).size == spread.size  # a random function: about 2000 collisions
    n = 1 << 22
    h = O.train_hash(np.arange(n, dtype=np.uint32), seed)
    assert np.unique(h).size == n
    bits = (h[:, None] >> np.arange(32, dtype=np.uint32)) & 1
    ones = bits.sum(0)
    assert np.all(np.abs(ones - n / 2) <= 6 * np.sqrt(n) / 2), ones
    counts = np.bincount(h >> 22, minlength=1024)
    chi2 = float(((counts - n / 1024) ** 2 / (n / 1024)).sum())
    assert chi2 < 1023 + 6 * np.sqrt(2 * 1023), chi2
    u = h.astype(np.float64) / 2 ** 32
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 6 / np.sqrt(n)


def test_the_bijection_check_fails_a_broken_hash():
    """The same check on a copy with an even LCG multiplier (one plausible typo)
    finds collisions: the check has the power to fail."""
    v = _spread_indices().astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    state = (v * np.uint64(747796404) + np.uint64(2891336453)) & m
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & m
    h = (word >> np.uint64(22)) ^ word
    assert np.unique(h).size < v.size


def test_seeds_give_independent_streams():
    i = np.arange(1 << 20, dtype=np.uint32)
    a, b = O.train_uniform(i, 9), O.train_uniform(i, 10)
    assert abs(np.corrcoef(a, b)[0, 1]) < 6 / np.sqrt(i.size)
    assert np.count_nonzero(a == b) < 64
    # float(hash) rounds to 24 bits, so hashes within 128 of 2^32 give exactly
    # 1.0, as the shader's float(word) / 4294967296.0 does: the range is [0, 1].
    assert np.all((a >= 0) & (a <= 1))


@pytest.mark.parametrize("p", [0.1, 0.5, 0.9])
def test_dropout_share_is_the_probability(p):
    n = 1 << 18
    for seed in (1, 9):
        dropped = int(np.count_nonzero(O.train_uniform(np.arange(n, dtype=np.uint32), seed) < np.float32(p)))
        assert _binomial_ok(dropped, p, n), (seed, dropped / n)


# The published constants of the hash (Jarzynski and Olano, "Hash Functions for
# GPU Rendering", JCGT 9(3), 2020, the "pcg" entry): one LCG step of PCG's
# 32-bit generator, multiplier 747796405 and increment 2891336453, then the
# RXS M XS output permutation with multiplier 277803737; the reference mixes
# the seed into the index first with 0x9E3779B9 (train_types.glsl:156-160).
_LCG_MUL, _LCG_ADD, _RXS_MUL, _SEED_MUL = 747796405, 2891336453, 277803737, 0x9E3779B9


def _unhash(h, seed):
    """The hash undone step by step with the published constants: xorshift 22
    undone, the multiplier's inverse modulo 2^32, the data-dependent xorshift
    undone (its shift is in the top four bits, which it leaves alone), the LCG
    step undone, the seed's mix undone. Nothing here restates the forward hash."""
    m = (1 << 32) - 1
    out = np.empty(h.size, np.uint32)
    inv_rxs, inv_lcg = pow(_RXS_MUL, -1, 1 << 32), pow(_LCG_MUL, -1, 1 << 32)
    for k, y in enumerate(h.tolist()):
        w = y ^ (y >> 22)
        s = (w * inv_rxs) & m
        r = (s >> 28) + 4
        state, shifted = s, s >> r
        while shifted:
            state ^= shifted
            shifted >>= r
        index = ((state - _LCG_ADD) * inv_lcg) & m
        out[k] = index ^ ((seed * _SEED_MUL) & m)
    return out


@pytest.mark.parametrize("seed", [0, 7, 0xFFFFFFFF])
def test_hash_is_the_published_pcg_permutation(seed):
    """Known answers by inversion: every hash of 20000 indices (spread over the
    32-bit range, with 0 and the largest) leads back to its index through the
    published constants, so the oracle's hash is exactly that function."""
    rng = np.random.default_rng(5)
    idx = np.concatenate([np.array([0, 1, 2, 0xFFFFFFFF], np.uint32), rng.integers(0, 1 << 32, 20000, dtype=np.uint64).astype(np.uint32)])
    np.testing.assert_array_equal(_unhash(O.train_hash(idx, seed), seed), idx)


def test_the_inversion_fails_a_hash_with_another_constant():
    """The inversion has the power to fail: a hash whose output multiplier is
    off by two does not lead back to its indices."""
    v = np.arange(1000, dtype=np.uint64)
    m = np.uint64(0xFFFFFFFF)
    state = (v * np.uint64(_LCG_MUL) + np.uint64(_LCG_ADD)) & m
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(_RXS_MUL + 2)) & m
    h = ((word >> np.uint64(22)) ^ word).astype(np.uint32)
    assert np.count_nonzero(_unhash(h, 0) != v.astype(np.uint32)) > 990
