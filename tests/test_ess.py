"""Effective sample size estimator (Geyer initial monotone sequence)."""
import numpy as np

from massivemds_amd import datagen
from massivemds_amd.hmc import effective_sample_size


def _ar1(rho, n, seed):
    e = datagen.normals(seed, 1, n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    return x


def test_iid():
    n = 20000
    ess = effective_sample_size(datagen.normals(3, 1, n))
    assert 0.85 * n < ess < 1.15 * n


def test_ar1_closed_form():
    # AR(1): integrated autocorrelation time (1 + rho) / (1 - rho)
    n = 40000
    for rho in (0.5, 0.9):
        ess = effective_sample_size(_ar1(rho, n, 11))
        expected = n * (1 - rho) / (1 + rho)
        assert 0.8 * expected < ess < 1.2 * expected, (rho, ess, expected)


def test_constant_series():
    assert np.isnan(effective_sample_size(np.ones(100)))
