"""GPU tests of the batched sampler (mds_hip_batch_hmc_run): many chains
in one launch per leapfrog step -- two chains sharing one dataset's
observations and one chain of another dataset of its own N and P -- each
equal, bit for bit, to the single-chain sampler (mds_hip_hmc_run) with
seed + k on the same data: the same kernels' arithmetic, reductions in the
same order, the same random streams.  Covers the precision's Metropolis
step (truncation) and its Gibbs draw (no truncation), the trait
precision's Wishart draws, step-size adaptation, the adapted width of the
precision's proposal, graph and direct launches, FP64 and FP32."""
import numpy as np
import pytest

from massivemds_amd import datagen, hmc
from massivemds_amd.batch import MultiDimensionalScalingBatch
from massivemds_amd.engine import Flags, MdsError, MultiDimensionalScaling

pytestmark = pytest.mark.gpu


def _datasets():
    da = datagen.euclid10(300, 51, 0.1, 0.02)
    db = datagen.euclid10(170, 52, 0.1, 0.02)
    return da, db


@pytest.mark.parametrize("truncated,fp32,graph,adapt_width", [(True, False, True, True), (False, False, True, False),
                                                              (True, True, False, False)])
def test_chains_equal_single_chain_runs(truncated, fp32, graph, adapt_width):
    print(f"batched hmc truncated={truncated} fp32={fp32} graph={graph} adapt_width={adapt_width}", flush=True)
    da, db = _datasets()
    flags = Flags.OPENCL | (Flags.LEFT_TRUNCATION if truncated else 0) | (Flags.FLOAT if fp32 else 0)
    dims, ns, taus = [2, 2, 3], [300, 300, 170], [1.5, 1.5, 0.8]
    data = [da, da, db]
    b = MultiDimensionalScalingBatch(dims, ns, flags, 0)
    b.setPairwiseData(0, da)
    b.sharePairwiseData(1, 0)
    b.setPairwiseData(2, db)
    for k, t in enumerate(taus):
        b.setParameters(k, [t])
    x0s = [datagen.locations(n, p, 60 + k) for k, (n, p) in enumerate(zip(ns, dims))]
    kw = dict(iterations=40, burnIn=15, leapfrogSteps=6, trajectory=0.06, learnPrecision=1, learnTraitPrecision=1,
              traitInvWeight=2, precisionProposalWidth=0.05, adaptStepSize=1, targetAccept=0.7,
              useGraph=int(graph), adaptPrecisionWidth=int(adapt_width))
    traces, finals, stats = b.run_hmc(x0s, hmc.default_options(seed=100, **kw))
    for k in range(3):
        e = MultiDimensionalScaling(dims[k], ns[k], flags, 0)
        e.setPairwiseData(data[k])
        e.setParameters([taus[k]])
        trace, final, st = hmc.run_hmc(e, x0s[k], hmc.default_options(seed=100 + k, **kw))
        e.close()
        np.testing.assert_array_equal(traces[k], trace)
        np.testing.assert_array_equal(finals[k], final)
        assert stats[k].acceptRate == st.acceptRate and stats[k].stepSize == st.stepSize
        assert stats[k].finalPrecision == st.finalPrecision
        assert stats[k].precisionProposalWidth == st.precisionProposalWidth
        assert np.all(np.isfinite(trace))
    # the chains of the shared dataset differ (own streams), the batch keeps each chain's precision
    assert not np.array_equal(traces[0], traces[1])
    b.evaluate()
    assert np.all(np.isfinite(b.sums()))
    b.close()


def test_sharing_and_argument_errors():
    da, db = _datasets()
    b = MultiDimensionalScalingBatch([2, 2, 3], [300, 300, 170], Flags.OPENCL | Flags.LEFT_TRUNCATION, 0)
    with pytest.raises(MdsError):
        b.sharePairwiseData(1, 0)  # the source has no data yet
    b.setPairwiseData(0, da)
    with pytest.raises(MdsError):
        b.sharePairwiseData(2, 0)  # another N
    with pytest.raises(MdsError):
        b.sharePairwiseData(0, 0)
    b.sharePairwiseData(1, 0)
    b.setPairwiseData(2, db)
    for k in range(3):
        b.setParameters(k, [1.0])
    x0s = [datagen.locations(n, p, 70 + k) for k, (n, p) in enumerate(zip([300, 300, 170], [2, 2, 3]))]
    b.updateLocations(0, x0s[0])
    b.updateLocations(1, x0s[0])
    b.updateLocations(2, x0s[2])
    b.evaluate()
    s = b.sums()
    assert s[0] == s[1]  # shared observations, same locations
    with pytest.raises(MdsError):
        b.run_hmc(x0s, hmc.default_options(iterations=5, burnIn=0, useSurrogate=1))
    with pytest.raises(MdsError):
        b.run_hmc(x0s, hmc.default_options(iterations=5, burnIn=5))
    # a valid run after the refusals; then data set on the sharing entry end the sharing
    traces, finals, stats = b.run_hmc(x0s, hmc.default_options(iterations=6, burnIn=2, leapfrogSteps=3,
                                                                trajectory=0.03))
    assert all(np.all(np.isfinite(t)) for t in traces)
    b.setPairwiseData(1, da * 1.1)
    for k in range(3):
        b.setParameters(k, [1.0])
    b.updateLocations(0, x0s[0])
    b.updateLocations(1, x0s[0])
    b.evaluate()
    s = b.sums()
    assert s[0] != s[1]
    b.close()
