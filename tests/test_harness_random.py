"""The exposure's random draws (optika_amd/csrc/optk_random.h), built for the
host by the test harness:

* Philox4x32-10 against the known-answer vectors published with the
  Random123 library (Salmon et al. 2011, kat_vectors: zero counter and key,
  all-ones, and the pi-digits case);
* Poisson draws against the Poisson law: sample mean and variance within
  five standard errors, and a chi-square test of the histogram against the
  exact probabilities, for means on both sides of the method switch at 10;
* Gaussian read noise: zero mean, the requested rms.
CPU only."""

import numpy as np
import pytest
import scipy.stats

from harness import expose, philox


@pytest.mark.parametrize("ctr,key,want", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, want):
    assert philox(ctr, key) == list(want)


@pytest.mark.parametrize("mean", [0.3, 3.0, 9.99, 10.0, 37.5, 1000.0, 2.5e6])
def test_poisson_law(mean):
    n = 200_000
    x = expose(np.full(n, mean), seed=12345)
    assert np.all(x == np.floor(x)) and np.all(x >= 0)
    se = np.sqrt(mean / n)
    assert abs(x.mean() - mean) < 5 * se, (x.mean(), mean)
    # variance of the sample variance of a Poisson: (mean + 2 mean^2 (n/(n-1))) / n
    sev = np.sqrt((mean + 2 * mean * mean) / n)
    assert abs(x.var() - mean) < 5 * sev, (x.var(), mean)
    lo, hi = scipy.stats.poisson.ppf([1e-4, 1 - 1e-4], mean)
    edges = np.arange(lo, hi + 2) if hi - lo < 200 else np.linspace(lo, hi + 1, 101).round()
    edges = np.unique(edges)
    obs, _ = np.histogram(x, bins=edges)
    p = np.diff(scipy.stats.poisson.cdf(edges - 1, mean))
    keep = p * n > 20
    obs, p = obs[keep], p[keep]
    exp_ = p / p.sum() * obs.sum()
    chi2 = ((obs - exp_) ** 2 / exp_).sum()
    assert scipy.stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, chi2


def test_zero_mean_and_read_noise():
    x = expose(np.zeros(1000), seed=1)
    assert np.all(x == 0)
    g = expose(np.zeros(200_000), seed=2, read_noise=3.0)
    assert abs(g.mean()) < 5 * 3.0 / np.sqrt(g.size)
    assert abs(g.std() - 3.0) < 0.02
    # draws depend on the pixel index and the seed only
    a, b = expose(np.full(10, 5.0), seed=7), expose(np.full(10, 5.0), seed=7)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, expose(np.full(10, 5.0), seed=8))
