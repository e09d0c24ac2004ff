"""Device-resident HMC sampler (counterpart of R/Untitled.R:477-729)."""
import ctypes
import math

import numpy as np
import pytest

from massivemds_amd import datagen, hmc
from massivemds_amd.engine import Flags, MultiDimensionalScaling, load_library

pytestmark = pytest.mark.gpu


def _options(**kw):
    opt = hmc.HmcOptions()
    hmc._bind(load_library()).mds_hip_hmc_default_options(ctypes.byref(opt))
    for k, v in kw.items():
        setattr(opt, k, v)
    return opt


def _engine(delta, p, tau, truncated=True):
    n = delta.shape[0]
    e = MultiDimensionalScaling(p, n, Flags.OPENCL | (Flags.LEFT_TRUNCATION if truncated else 0), 0)
    e.setPairwiseData(delta)
    e.setParameters([tau])
    return e


def _functionals(x):
    n = x.shape[0]
    return np.array([np.linalg.norm(x[0] - x[1]), np.linalg.norm(x[0]), np.linalg.norm(x[n - 1])])


def test_hmc_matches_random_walk_metropolis(oracle):
    """Same target, independent sampler: random-walk Metropolis on the oracle's
    potential.  Posterior means of invariant functionals must agree."""
    n, p, tau = 6, 2, 1.5
    delta = datagen.euclid10(n, seed=21) * 0.4
    x0 = datagen.locations(n, p, seed=22)
    e = _engine(delta, p, tau)
    opt = _options(iterations=6000, burnIn=500, leapfrogSteps=10, trajectory=0.6, learnPrecision=0,
                   learnTraitPrecision=0, seed=5)
    trace, _, stats = hmc.run_hmc(e, x0, opt)
    assert 0.5 < stats.acceptRate <= 1.0
    h = trace[:, [3, 8, 10]]  # d_0_1, |x_0|, |x_{N-1}|

    pairs = n * (n - 1) / 2

    def potential(x):
        s = oracle.sum_of_increments(delta, x, tau, True)
        ll = 0.5 * (math.log(tau) - math.log(2 * math.pi)) * pairs - s
        return -(ll - 0.5 * float(np.sum(x * x)))

    rng = np.random.default_rng(7)
    x = x0.copy()
    u = potential(x)
    samples = []
    steps = 120000
    prop = rng.normal(size=(steps, n, p)) * 0.12
    logu = np.log(rng.uniform(size=steps))
    for t in range(steps):
        y = x + prop[t]
        uy = potential(y)
        if u - uy > logu[t]:
            x, u = y, uy
        if t >= 5000 and t % 5 == 0:
            samples.append(_functionals(x))
    r = np.array(samples)
    for k in range(3):
        se_h = h[:, k].std() / math.sqrt(hmc.effective_sample_size(h[:, k]))
        se_r = r[:, k].std() / math.sqrt(hmc.effective_sample_size(r[:, k]))
        assert abs(h[:, k].mean() - r[:, k].mean()) <= 5 * math.hypot(se_h, se_r), (
            k, h[:, k].mean(), r[:, k].mean(), se_h, se_r)
    e.close()


def test_graph_and_direct_launches_identical():
    n, p = 200, 2
    delta = datagen.euclid10(n, seed=3, missing=0.02)
    x0 = datagen.locations(n, p, seed=4)
    out = []
    for graph in (1, 0):
        e = _engine(delta, p, 1.0)
        opt = _options(iterations=12, burnIn=2, leapfrogSteps=5, trajectory=0.05, useGraph=graph, seed=9)
        trace, final, stats = hmc.run_hmc(e, x0, opt)
        out.append((trace, final))
        e.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("truncation", [True, False])
def test_learning_smoke(truncation):
    n = 400
    delta = datagen.euclid10(n, seed=31, missing=0.02)
    res = hmc.hmcsampler(40, burnIn=5, data=delta, latentDimension=2, learnPrec=True, learnTraitPrec=True,
                         trajectory=0.05, mdsPrecision=1.0, truncation=truncation, seed=3)
    assert np.all(np.isfinite(res["trace"]))
    assert 0.0 <= res["acprat"] <= 1.0
    assert res["locations"].shape == (n, 2)
    prec = res["precision"]
    assert np.all(prec > 0)
    if not truncation:
        assert res["precision_acceptance"] == 1.0  # Gibbs
    assert res["trace"][:, 2].std() > 0  # trait precision moves


def test_argument_and_state_errors():
    """The sampler refuses to start without data or a precision, and rejects
    bad iteration counts, trajectories, trait-update weights, initial
    locations, a surrogate mode without a surrogate and a surrogate built for
    another N, each with a status and a message; the engine still samples
    afterwards."""
    from massivemds_amd.engine import MdsError
    from massivemds_amd.surrogate import Surrogate

    n, p = 40, 2
    delta = datagen.euclid10(n, seed=5)
    x0 = datagen.locations(n, p, seed=6)
    e = MultiDimensionalScaling(p, n, Flags.OPENCL | Flags.LEFT_TRUNCATION, 0)
    with pytest.raises(MdsError) as info:
        hmc.run_hmc(e, x0, _options(iterations=4, burnIn=1))
    assert info.value.status == 3 and "data" in str(info.value)
    e.setPairwiseData(delta)
    with pytest.raises(MdsError) as info:
        hmc.run_hmc(e, x0, _options(iterations=4, burnIn=1))
    assert info.value.status == 3 and "precision" in str(info.value)
    e.setParameters([1.5])
    bad = [dict(iterations=0, burnIn=0), dict(iterations=4, burnIn=4), dict(iterations=4, burnIn=1, leapfrogSteps=0),
           dict(iterations=4, burnIn=1, trajectory=0.0), dict(iterations=4, burnIn=1, traitInvWeight=0),
           dict(iterations=4, burnIn=1, useSurrogate=1)]
    for kw in bad:
        with pytest.raises(MdsError) as info:
            hmc.run_hmc(e, x0, _options(**kw))
        assert info.value.status == 1, kw
    with pytest.raises(MdsError) as info:
        hmc.run_hmc(e, x0[:-1], _options(iterations=4, burnIn=1))
    assert info.value.status == 1
    other = Surrogate(n + 1, p, 4)
    for kw in (dict(useSurrogate=1), dict(collectSnapshots=1)):
        with pytest.raises(MdsError) as info:
            hmc.run_hmc(e, x0, _options(iterations=4, burnIn=1, surrogate=other.handle, **kw))
        assert info.value.status == 1 and "another N" in str(info.value), kw
    other.close()
    trace, final, stats = hmc.run_hmc(e, x0, _options(iterations=4, burnIn=1, leapfrogSteps=5, trajectory=0.05))
    assert trace.shape == (3, len(hmc.TRACE_COLUMNS)) and np.all(np.isfinite(trace)) and np.all(np.isfinite(final))
    e.close()
