"""The replay stream: std::mt19937_64 + uniform_real_distribution<double>."""
import numpy as np


def test_mt19937_64_known_answer(oracle_mod):
    # C++ [rand.predef]: the 10000th consecutive invocation of a
    # default-constructed mt19937_64 (seed 5489) produces 9981545732273789042
    raw = oracle_mod.mt19937_64_raw(5489, 10000)
    assert int(raw[-1]) == 9981545732273789042


def test_uniform_is_the_top_53_bits_scaled(oracle_mod):
    raw = oracle_mod.mt19937_64_raw(1, 4096)
    u = oracle_mod.mt19937_64_uniform(1, 4096)
    # generate_canonical<double,53> over one 64-bit word: x / 2^64
    expect = raw.astype(np.float64) / 2.0 ** 64
    expect = np.where(expect >= 1.0, np.nextafter(1.0, 0.0), expect)
    assert np.array_equal(u, expect)
    assert u.min() >= 0 and u.max() < 1


def test_streams_differ_by_seed(oracle_mod):
    a = oracle_mod.mt19937_64_uniform(1, 100)
    b = oracle_mod.mt19937_64_uniform(2, 100)
    assert not np.array_equal(a, b)
