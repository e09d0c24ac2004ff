"""The synthetic observation generator: euclid10's distances equal the
direct expression sqrt(sum((z_i - z_j)^2)) (numpy's own summation) bit for
bit, for dimensions above and below numpy's eight-way pairwise block, ragged
row blocks included; and the noise / missing rules."""
import numpy as np
import pytest

from massivemds_amd import datagen


@pytest.mark.parametrize("n,dim,block", [(1, 10, 512), (7, 10, 3), (300, 10, 64), (129, 3, 50), (65, 12, 512)])
def test_distances_equal_direct_expression(n, dim, block):
    z = datagen.normals(5, 1, n * dim).reshape(n, dim)
    direct = np.sqrt(np.sum((z[:, None, :] - z[None, :, :]) ** 2, axis=2))
    assert np.array_equal(datagen._distances(z, block), direct)


def test_euclid10_rules():
    d = datagen.euclid10(200, seed=9, noise_sd=0.1, missing=0.05)
    assert np.all(np.diag(d) == 0.0)
    both = np.isnan(d)
    assert np.array_equal(both, both.T)
    assert 0.03 < both[np.triu_indices(200, 1)].mean() < 0.07
    ok = ~both
    assert np.array_equal(d[ok], d.T[ok]) and np.all(d[ok] >= 0.0)
