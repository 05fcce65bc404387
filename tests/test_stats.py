"""Bin statistics of the MC_Stat_Results files (host code, CPU only): the
MT19937 stream the bootstrap draws from, the mean and the two deviation
rules of get_statVectors (CF/aux_code.cpp:1050-1187)."""
import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H


def test_mt19937_known_answers():
    # the C++ standard fixes the 10000th output of mt19937 seeded 5489
    assert int(H.debug_mt19937(5489, 10000)[-1]) == 4123659995
    # numpy's legacy RandomState seeds MT19937 the same way (init_genrand)
    st = np.random.RandomState(20240).get_state()
    assert st[0] == "MT19937"
    ref = np.random.RandomState(20240)
    words = ref.randint(0, 2 ** 32, size=8, dtype=np.uint64)
    got = H.debug_mt19937(20240, 8).astype(np.uint64)
    # RandomState draws 32-bit words for a 2^32 range through a mask: raw outputs
    np.testing.assert_array_equal(got, words)
    # GSL's convention: seed 0 means 4357
    np.testing.assert_array_equal(H.debug_mt19937(0, 5), H.debug_mt19937(4357, 5))


def test_means_and_sample_deviation():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(40, 7)) * np.arange(1, 8)
    m, s = H.bin_statistics(d, bootstrap=False)
    np.testing.assert_allclose(m, d.mean(0), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(s, d.std(0, ddof=1), rtol=1e-13)
    mb, sb = H.bin_statistics(d, bootstrap=True, seed=11)
    np.testing.assert_allclose(mb, d.mean(0), rtol=1e-14, atol=1e-15)
    # bootstrap deviation of the mean: population sd / sqrt(n), within the
    # ~2.2 % spread of 1000 resamples (5 sigma)
    want = d.std(0, ddof=0) / np.sqrt(40)
    assert np.all(np.abs(sb / want - 1) < 0.12)
    # deterministic in the seed, and the seed matters
    np.testing.assert_array_equal(sb, H.bin_statistics(d, True, 11)[1])
    assert not np.array_equal(sb, H.bin_statistics(d, True, 12)[1])


def test_bootstrap_of_a_two_bin_indicator():
    # resampled means of (1, 0) are Binomial(2, 1/2) / 2: sd = sqrt(1/8)
    d = np.array([[1.0], [0.0]])
    _, sb = H.bin_statistics(d, bootstrap=True, seed=3)
    assert sb[0] == pytest.approx(np.sqrt(1 / 8), rel=0.08)
    # a constant column has no spread
    _, s0 = H.bin_statistics(np.full((10, 1), 2.5), bootstrap=True, seed=3)
    assert s0[0] == 0.0
    with pytest.raises(H.HSMDError):
        H.bin_statistics(np.zeros((1, 3)))
