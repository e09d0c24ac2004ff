"""The engine's read-backs through the C ABI: a call that runs a pass gets the
sum (and, when it reads it, the gradient) written by the reduce kernel into
pinned host memory; a call after a device-resident pass
(mds_hip_evaluate_async) copies them from the device instead.  Both routes
must give the same bits, whichever call comes first and whether or not the
gradient is predicted with the sum (src/NewMultiDimensionalScaling.hpp:193-203,
:281-302 for the values the calls return).
"""
import numpy as np
import pytest

from massivemds_amd import datagen
from massivemds_amd.engine import Flags, MultiDimensionalScaling

pytestmark = pytest.mark.gpu


def _engine(delta, x, p, tau, fp32):
    n = delta.shape[0]
    e = MultiDimensionalScaling(p, n, Flags.OPENCL | Flags.LEFT_TRUNCATION | (Flags.FLOAT if fp32 else 0), 0)
    e.setPairwiseData(delta)
    e.updateLocations(-1, x)
    e.setParameters([tau])
    return e


@pytest.mark.parametrize("n,p,fp32", [(300, 2, False), (257, 3, False), (300, 2, True), (130, 5, True)])
def test_kernel_written_and_copied_results_agree(oracle, n, p, fp32):
    print(f"readback n={n} p={p} fp32={fp32}", flush=True)
    tau = 1.7
    delta = datagen.euclid10(n, seed=n, noise_sd=0.1, missing=0.05)
    x = datagen.locations(n, p, seed=n + 1)

    # copied: a device-resident pass first, then the calls read its results
    a = _engine(delta, x, p, tau, fp32)
    a.evaluate_async(True)
    a.synchronize()
    s_copy = a.getSumOfIncrements()
    g_copy = a.getLogLikelihoodGradient()

    # written by the kernels: the sum alone, then the gradient (two passes)
    b = _engine(delta, x, p, tau, fp32)
    s_sum = b.getSumOfIncrements()
    g_after = b.getLogLikelihoodGradient()

    # the gradient first: it brings the sum with it
    c = _engine(delta, x, p, tau, fp32)
    g_first = c.getLogLikelihoodGradient()
    s_after = c.getSumOfIncrements()

    # the gradient predicted with the sum: after a few states whose gradient
    # was read, the sum call runs the fused pass and writes both
    d = _engine(delta, x, p, tau, fp32)
    for k in range(12):
        d.updateLocations(-1, datagen.locations(n, p, seed=1000 + k))
        d.getSumOfIncrements()
        d.getLogLikelihoodGradient()
    d.updateLocations(-1, x)
    s_pred = d.getSumOfIncrements()
    g_pred = d.getLogLikelihoodGradient()

    for s in (s_sum, s_after, s_pred):
        assert s == s_copy
    for g in (g_after, g_first, g_pred):
        assert np.array_equal(g, g_copy)
    ref = oracle.sum_of_increments(delta, x, tau, True)
    assert abs(s_copy - ref) <= (2e-6 if fp32 else 1e-12) * abs(ref)
    for e in (a, b, c, d):
        e.close()
