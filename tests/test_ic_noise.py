"""Host half of the counter-based noise stream (CPU): known-answer values of SplitMix64 and the
statistics of the Box-Muller field."""
import numpy as np

from pbh_amd import ic


def test_splitmix64_known_answers():
    # first outputs of the published SplitMix64 generator seeded with 0: the finaliser applied to
    # the Weyl sequence 0x9e3779b97f4a7c15 * k (state after k increments, k = 1, 2, 3)
    seq = [int(v) for v in ic.splitmix64(np.array([0, 0x9E3779B97F4A7C15, (2 * 0x9E3779B97F4A7C15) % 2**64],
                                                  dtype=np.uint64))]
    assert seq == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_white_noise_statistics_and_shape_independence():
    w = ic.white_noise(7, (32, 32, 32))
    assert abs(w.mean()) < 4 / np.sqrt(w.size) and abs(w.std() - 1) < 0.01  # 4 sigma
    assert np.array_equal(w.reshape(-1)[:100], ic.white_noise(7, (100,)))
    assert not np.array_equal(w, ic.white_noise(8, (32, 32, 32)))
