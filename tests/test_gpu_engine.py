"""GPU parity of the HIP engine against the oracle, through the C ABI.

Bar (the FP64 target of the engine): log-likelihood sum within 1e-12
relative, gradient within 1e-12 of max|g| (norm-relative; entries near zero
are sums of signed terms), against the CPU restatement of the reference
serial engine.  The FP32 engine is held to 1e-6 relative on the
log-likelihood against the FP64 oracle.
"""
import math

import numpy as np
import pytest

from massivemds_amd import datagen, mds
from massivemds_amd.engine import Flags, MdsError, MultiDimensionalScaling
from tests.golden_data import beast_kat

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def _flags(truncated, fp32=False):
    return Flags.OPENCL | (Flags.LEFT_TRUNCATION if truncated else 0) | (Flags.FLOAT if fp32 else 0)


def _data(n, p, pattern, seed=5):
    delta = datagen.euclid10(n, seed=seed, noise_sd=0.2, missing=0.05 if pattern == "sym5" else 0.0)
    if pattern == "row" and n > 1:
        delta[1, :] = np.nan
        delta[:, 1] = np.nan
    if pattern == "asym":
        u = datagen.normals(seed, 9, n * n).reshape(n, n)
        delta = np.abs(delta + 0.1 * u)
        drop = datagen.uniforms(seed, 10, n * n).reshape(n, n) < 0.07
        delta[drop] = np.nan
        np.fill_diagonal(delta, 0.0)
    x = datagen.locations(n, p, seed=seed + 100)
    return delta, x


def _gpu(delta, x, tau, truncated, fp32=False):
    n, p = x.shape
    e = MultiDimensionalScaling(p, n, _flags(truncated, fp32), 0)
    e.setPairwiseData(delta)
    e.updateLocations(-1, x)
    e.setParameters([tau])
    s = e.getSumOfIncrements()
    g = e.getLogLikelihoodGradient().reshape(n, p)
    return e, s, g


def _check(s, g, s_ref, g_ref, rtol):
    assert abs(s - s_ref) <= rtol * max(abs(s_ref), 1e-300), (s, s_ref, abs(s - s_ref) / abs(s_ref))
    gmax = np.max(np.abs(g_ref)) if g_ref.size else 0.0
    err = np.max(np.abs(g - g_ref)) if g_ref.size else 0.0
    assert err <= rtol * max(gmax, 1e-300), (err, gmax)


CASES = [
    # n, p, pattern
    (2, 2, "none"), (3, 1, "none"), (17, 2, "sym5"), (63, 3, "none"), (64, 2, "sym5"),
    (65, 2, "row"), (130, 5, "sym5"), (200, 2, "asym"), (257, 10, "sym5"), (129, 16, "asym"),
    (300, 4, "none"), (100, 6, "row"), (90, 8, "asym"), (70, 12, "none"), (1000, 2, "sym5"),
    (150, 5, "asym"), (300, 3, "row"),
]


@pytest.mark.parametrize("n,p,pattern", CASES)
@pytest.mark.parametrize("truncated", [True, False])
def test_fp64_parity(oracle, n, p, pattern, truncated):
    tau = {2: 2.0, 3: 0.5}.get(p, 1.0)
    delta, x = _data(n, p, pattern)
    e, s, g = _gpu(delta, x, tau, truncated)
    _check(s, g, oracle.sum_of_increments(delta, x, tau, truncated),
           oracle.gradient(delta, x, tau, truncated), RTOL64)
    sym, nbytes, tiles = e.layout()
    assert sym == (pattern != "asym")


@pytest.mark.parametrize("n,p,pattern", [(257, 2, "sym5"), (1000, 2, "sym5"), (130, 5, "asym"), (200, 10, "none"),
                                         (70, 12, "row"), (129, 16, "asym")])
@pytest.mark.parametrize("truncated", [True, False])
def test_fp32_error_against_fp64_oracle(oracle, n, p, pattern, truncated):
    tau = 2.0
    delta, x = _data(n, p, pattern)
    _, s, g = _gpu(delta, x, tau, truncated, fp32=True)
    s_ref = oracle.sum_of_increments(delta, x, tau, truncated)
    g_ref = oracle.gradient(delta, x, tau, truncated)
    ll = oracle.log_likelihood(s, n, tau)
    ll_ref = oracle.log_likelihood(s_ref, n, tau)
    print(f"fp32 n={n} p={p} {pattern} T={truncated}: rel(S)={abs(s - s_ref) / abs(s_ref):.2e} "
          f"rel(ll)={abs(ll - ll_ref) / abs(ll_ref):.2e} "
          f"grad={np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref)):.2e}")
    assert abs(ll - ll_ref) <= 1e-6 * abs(ll_ref)
    assert abs(s - s_ref) <= 1e-6 * abs(s_ref)
    assert np.max(np.abs(g - g_ref)) <= 1e-5 * np.max(np.abs(g_ref))


def test_edge_single_location(oracle):
    delta = np.zeros((1, 1))
    x = np.array([[0.3, -0.2]])
    for t in (True, False):
        _, s, g = _gpu(delta, x, 1.0, t)
        assert s == 0.0 and not np.any(g)


def test_edge_all_missing():
    n = 70
    delta = np.full((n, n), np.nan)
    x = datagen.locations(n, 3)
    _, s, g = _gpu(delta, x, 2.0, True)
    assert s == 0.0 and not np.any(g)


def test_beast_known_answer_gpu():
    k = beast_kat()
    for fp32, tol in ((False, 2e-9), (True, 1e-6 * abs(k["loglik"]))):
        eng = mds.createEngine(2, len(k["x"]), truncation=False, gpu=1, single=fp32)
        mds.setPairwiseData(eng, k["delta"])
        mds.updateLocations(eng, k["x"])
        mds.setPrecision(eng, k["precision"])
        ll = mds.getLogLikelihood(eng)
        assert abs(ll - k["loglik"]) < tol, (fp32, ll, k["loglik"])


def test_deterministic_bitwise():
    delta, x = _data(700, 2, "sym5")
    e, s, g = _gpu(delta, x, 2.0, True)
    for _ in range(3):
        e.makeDirty()
        e.updateLocations(-1, x)
        assert e.getSumOfIncrements() == s
        np.testing.assert_array_equal(e.getLogLikelihoodGradient().reshape(g.shape), g)


def test_caching_semantics(oracle):
    """src/NewMultiDimensionalScaling.hpp:134-302 state rules through the C ABI."""
    n, p = 150, 2
    delta, x = _data(n, p, "sym5")
    x2 = x + 0.1
    # untruncated: setParameters keeps the cached sum and rescales it (:198-202, :262-274)
    e, s1, _ = _gpu(delta, x, 1.0, False)
    e.setParameters([3.0])
    assert e.getSumOfIncrements() == 3.0 * s1
    assert abs(e.getSumOfIncrements() - oracle.sum_of_increments(delta, x, 3.0, False)) <= 1e-12 * 3 * s1
    # truncated: setParameters invalidates
    e, s1, g1 = _gpu(delta, x, 1.0, True)
    e.setParameters([2.0])
    assert abs(e.getSumOfIncrements() - oracle.sum_of_increments(delta, x, 2.0, True)) <= 1e-12 * abs(s1)
    e.setParameters([1.0])
    assert e.getSumOfIncrements() == s1  # recomputed, bitwise identical
    # store / update / restore (:205-255); storeState keeps the cached member
    e.storeState()
    e.updateLocations(-1, x2)
    s2 = e.getSumOfIncrements()
    assert abs(s2 - oracle.sum_of_increments(delta, x2, 1.0, True)) <= 1e-12 * abs(s2)
    e.restoreState()
    assert e.getSumOfIncrements() == s1
    g = e.getLogLikelihoodGradient().reshape(n, p)
    np.testing.assert_array_equal(g, g1)
    # single-row update equals a full update with that row changed
    row = np.array([0.5, -1.25])
    e.updateLocations(7, row)
    x3 = x.copy()
    x3[7] = row
    _check(e.getSumOfIncrements(), e.getLogLikelihoodGradient().reshape(n, p),
           oracle.sum_of_increments(delta, x3, 1.0, True), oracle.gradient(delta, x3, 1.0, True), RTOL64)
    e.acceptState()
    e.makeDirty()
    assert e.getInternalDimension() == p


def test_gradient_reuse_and_prediction():
    """Callers asking for the sum and then the gradient of each state get
    one fused pass per state once more than 0.3 of the states have had their
    gradient asked for (the fourth state here), with results bitwise equal
    to separate evaluations; repeated gradient calls of one state launch
    nothing; after storeState / move / restoreState the restored state's
    gradient comes back without a pass (the reference benchmark's loop,
    src/benchmark.cpp:254-307); sums alone then cost one pass each and a
    later gradient is computed, not stale."""
    n, p = 300, 2
    delta, x = _data(n, p, "sym5")
    xs = [x + 0.05 * k for k in range(8)]
    ref = []
    for xk in xs:  # separate engines: a likelihood pass, then a fused pass
        e, s, g = _gpu(delta, xk, 2.0, True)
        ref.append((s, g))
        e.close()
    e = MultiDimensionalScaling(p, n, _flags(True), 0)
    e.setPairwiseData(delta)
    e.setParameters([2.0])
    f0, _ = e.counters()
    for k, xk in enumerate(xs):
        e.updateLocations(-1, xk)
        s = e.getSumOfIncrements()
        g = e.getLogLikelihoodGradient().reshape(n, p)
        assert s == ref[k][0]
        np.testing.assert_array_equal(g, ref[k][1])
        np.testing.assert_array_equal(e.getLogLikelihoodGradient().reshape(n, p), g)
    f1, _ = e.counters()
    assert f1 - f0 == len(xs) + 3  # two passes for each of the first three states
    e.updateLocations(-1, xs[0])
    np.testing.assert_array_equal(e.getLogLikelihoodGradient().reshape(n, p), ref[0][1])
    f1, _ = e.counters()
    for k in range(1, len(xs)):  # store, move, sum, reject, gradient of the restored state
        e.storeState()
        e.updateLocations(-1, xs[k])
        assert e.getSumOfIncrements() == ref[k][0]
        e.restoreState()
        assert e.getSumOfIncrements() == ref[0][0]
        np.testing.assert_array_equal(e.getLogLikelihoodGradient().reshape(n, p), ref[0][1])
    f2, _ = e.counters()
    assert f2 - f1 == len(xs) - 1  # the moved states' sums only
    e.storeState()  # a new precision drops the stored gradient
    e.updateLocations(-1, xs[1])
    e.setParameters([2.0])
    e.restoreState()
    np.testing.assert_array_equal(e.getLogLikelihoodGradient().reshape(n, p), ref[0][1])
    f3, _ = e.counters()
    assert f3 - f2 == 1
    for k, xk in enumerate(xs):  # sums alone
        e.updateLocations(-1, xk)
        assert e.getSumOfIncrements() == ref[k][0]
    f4, _ = e.counters()
    assert f4 - f3 == len(xs)
    e.updateLocations(-1, xs[0])  # and a gradient afterwards is computed, not stale
    np.testing.assert_array_equal(e.getLogLikelihoodGradient().reshape(n, p), ref[0][1])
    e.close()


def test_argument_errors():
    delta, x = _data(40, 2, "none")
    e = MultiDimensionalScaling(2, 40, _flags(True), 0)
    with pytest.raises(MdsError) as info:
        e.getSumOfIncrements()
    assert info.value.status == 3
    with pytest.raises(MdsError):
        e.setPairwiseData(delta[:10])
    e.setPairwiseData(delta)
    with pytest.raises(MdsError):
        e.updateLocations(-1, x[:5])
    with pytest.raises(MdsError):
        e.updateLocations(40, x[0])
    with pytest.raises(MdsError):
        e.setParameters([0.0])
    with pytest.raises(MdsError):
        e.setParameters([1.0, 2.0])
    e.updateLocations(-1, x)
    e.setParameters([1.0])
    with pytest.raises(MdsError):
        e.getLogLikelihoodGradient(10)
    assert np.isfinite(e.getSumOfIncrements())


def test_reference_benchmark_loop(oracle):
    """Config N=1000, P=2, tau=1, truncation: the reference benchmark's
    store -> update -> likelihood -> restore/accept -> gradient loop
    (src/benchmark.cpp:250-309) with fresh locations each iteration, every
    iteration checked against the oracle."""
    n, p = 1000, 2
    delta = datagen.euclid10(n)
    e = MultiDimensionalScaling(p, n, _flags(True), 0)
    e.setPairwiseData(delta)
    x = datagen.locations(n, p)
    e.updateLocations(-1, x)
    e.setParameters([1.0])
    current = x
    restore = datagen.uniforms(42, 1, 10) < 0.75
    for it in range(10):
        e.storeState()
        proposal = datagen.locations(n, p, seed=1000 + it)
        e.updateLocations(-1, proposal)
        s = e.getSumOfIncrements()
        assert abs(s - oracle.sum_of_increments(delta, proposal, 1.0, True)) <= RTOL64 * abs(s)
        if restore[it]:
            e.restoreState()
        else:
            e.acceptState()
            current = proposal
        g = e.getLogLikelihoodGradient().reshape(n, p)
        g_ref = oracle.gradient(delta, current, 1.0, True)
        assert np.max(np.abs(g - g_ref)) <= RTOL64 * np.max(np.abs(g_ref))


@pytest.mark.parametrize("p,pattern", [(2, "asym"), (2, "sym5"), (1, "row")])
@pytest.mark.parametrize("truncated", [True, False])
def test_eight_wavefront_launch(oracle, p, pattern, truncated):
    """FP64 walks with P <= 2 run as four-wavefront workgroups up to 1024
    tiles (N <= 2,816) and as eight-wavefront ones above: every variant of
    the larger launch (truncated or not, symmetric or asymmetric
    observations, a missing row) against the oracle at N = 3,000."""
    n, tau = 3000, 1.7
    print(f"eight-wavefront launch n={n} p={p} pattern={pattern} truncated={truncated}", flush=True)
    delta, x = _data(n, p, pattern)
    e, s, g = _gpu(delta, x, tau, truncated)
    e.close()
    _check(s, g, oracle.sum_of_increments(delta, x, tau, truncated), oracle.gradient(delta, x, tau, truncated),
           RTOL64)


def test_full_size_n10000(oracle):
    """N = 10,000, P = 2, tau = 2, 5% missing, noise 0.1: FP64 against the
    multi-core oracle (validated against the serial one at small N), FP32
    error against FP64, and the size-independent invariants."""
    n, p, tau = 10000, 2, 2.0
    delta = datagen.euclid10(n, noise_sd=0.1, missing=0.05)
    x = datagen.locations(n, p)
    e, s, g = _gpu(delta, x, tau, True)
    s_ref = oracle.simd_sum_of_increments(delta, x, tau, True)
    g_ref = oracle.simd_gradient(delta, x, tau, True)
    _check(s, g, s_ref, g_ref, RTOL64)
    mag = np.abs(g).sum()
    assert np.max(np.abs(g.sum(axis=0))) <= 1e-13 * mag
    m = x.T @ g
    assert abs(m[0, 1] - m[1, 0]) <= 1e-12 * np.abs(x).max() * mag
    e32, s32, g32 = _gpu(delta, x, tau, True, fp32=True)
    ll, ll32 = oracle.log_likelihood(s_ref, n, tau), oracle.log_likelihood(s32, n, tau)
    print(f"N=10000 fp32: rel(S)={abs(s32 - s_ref) / abs(s_ref):.2e} rel(ll)={abs(ll32 - ll) / abs(ll):.2e} "
          f"grad={np.max(np.abs(g32 - g_ref)) / np.max(np.abs(g_ref)):.2e}")
    assert abs(ll32 - ll) <= 1e-6 * abs(ll)
