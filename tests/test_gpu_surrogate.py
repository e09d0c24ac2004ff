"""GPU tests of the per-location MLP surrogate (massivemds_amd/csrc/mds_mlp.hip)
through the C ABI: forward / backward kernels against the CPU oracle
(oracle/mlp_oracle.py, itself pinned by finite differences), training on
exact gradient snapshots, the explained gradient variance the trained
network gains over no network (against the per-location ceiling of the
Gaussian approximation), and the surrogate-trajectory sampler's posterior
against plain HMC.  The network has no counterpart in the reference."""
import numpy as np
import pytest

from massivemds_amd import datagen, hmc
from massivemds_amd.engine import Flags, MultiDimensionalScaling
from massivemds_amd.surrogate import (Surrogate, map_locations, per_location_ceiling, plugin_precision,
                                      residual_hessian, split_parameters)
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _network(n, p, seed):
    sur = Surrogate(n, p, 4, seed=seed)
    rng = np.random.default_rng(seed)
    params = sur.get_parameters()
    params = (params + 0.05 * rng.standard_normal(params.size)).astype(np.float32)
    sur.set_parameters(params)
    xc = rng.standard_normal((n, p))
    gc = rng.standard_normal((n, p))
    sx = 0.5 + rng.random(p)
    sg = 0.5 + rng.random(p)
    sur.set_normalisation(xc, gc, sx, sg)
    oracle_params = [w.astype(np.float64) for w in split_parameters(params, n)]
    pad = lambda v: np.concatenate([np.asarray(v, np.float32), np.ones(16 - p, np.float32)]).astype(np.float64)
    return sur, oracle_params, xc.astype(np.float32).astype(np.float64), gc.astype(np.float32).astype(np.float64), \
        pad(sx), pad(sg)


@pytest.mark.parametrize("n,p", [(70, 3), (129, 5), (64, 16), (5, 1)])
def test_loss_gradient_matches_oracle(n, p):
    print(f"loss_gradient n={n} p={p}", flush=True)
    sur, params, xc, gc, sx, sg = _network(n, p, seed=n + p)
    x = datagen.normals(11, 1, n * p).reshape(n, p)
    g = 3.0 * datagen.normals(11, 2, n * p).reshape(n, p)
    loss, grad = sur.loss_gradient(x, g)
    loss_o, grads_o = mo.forward_backward(params, x, g, xc, gc, sx, sg, p, bf16=True)
    assert abs(loss - loss_o) <= 1e-4 * loss_o
    blocks = split_parameters(grad, n)
    names = ["W1", "b1", "W2", "b2", "W3", "b3", "E"]
    for name, gpu, ref in zip(names, blocks, grads_o):
        assert _rel(gpu, ref) <= 2e-3, (name, _rel(gpu, ref))
    # against the exact FP64 network the BF16 kernels agree to BF16 accuracy
    loss64, grads64 = mo.forward_backward(params, x, g, xc, gc, sx, sg, p, bf16=False)
    assert abs(loss - loss64) <= 2e-2 * loss64
    for name, gpu, ref in zip(names, blocks, grads64):
        assert _rel(gpu, ref) <= 5e-2, (name, _rel(gpu, ref))
    sur.close()


@pytest.mark.parametrize("n,p", [(70, 3), (300, 2)])
def test_predict_matches_oracle(n, p):
    sur, params, xc, gc, sx, sg = _network(n, p, seed=7 * n + p)
    x = xc + 0.7 * datagen.normals(12, 1, n * p).reshape(n, p)
    pred = sur.predict(x)
    ref = mo.predict(params, x, xc, gc, sx, sg, p, bf16=True)
    assert _rel(pred - gc, ref - gc) <= 2e-3
    sur.close()


def test_argument_errors():
    with pytest.raises(Exception):
        Surrogate(10, 17, 4)
    sur = Surrogate(10, 2, 2)
    with pytest.raises(Exception):
        sur.train(1, 1e-3)  # no snapshots
    x = np.zeros((10, 2))
    sur.add_snapshot(x, x)
    sur.add_snapshot(x, x)
    with pytest.raises(Exception):
        sur.add_snapshot(x, x)  # store full
    with pytest.raises(Exception):
        sur.predict(np.zeros(7))
    with pytest.raises(Exception):
        sur.set_normalisation(np.zeros((10, 2)), np.zeros((10, 2)), [1.0, 0.0], [1.0, 1.0])
    assert sur.snapshot_count() == 2
    sur.close()


def _problem(n, p, seed, truncation, precision=None, dim=10, noise=0.1):
    """Data (distances of N(0, I_dim) points with N(0, noise^2) noise, 2%
    missing), an engine and its posterior mode (plug-in precision if None)."""
    delta = datagen.euclid10(n, seed, noise, 0.02, dim=dim)
    flags = Flags.OPENCL | (Flags.LEFT_TRUNCATION if truncation else 0)
    eng = MultiDimensionalScaling(p, n, flags, 0)
    eng.setPairwiseData(delta)
    eng.setParameters([1.0 if precision is None else precision])
    x = map_locations(eng, datagen.normals(seed, 3, n * p).reshape(n, p))
    if precision is None:
        precision = plugin_precision(delta, x)
        eng.setParameters([precision])
        x = map_locations(eng, x)
    return delta, eng, x, precision


def test_training_gain_against_ceiling():
    """Snapshots from plain HMC at stationarity on the exact target; the
    trained network's explained variance of held-out log-posterior gradient
    rows against (a) no network (the best constant per location, fitted on
    the held-out rows themselves: 0) and (b) the best any per-location model
    can do under the Gaussian approximation at the mode."""
    n, p, L = 300, 2, 10
    delta, eng, x, tau = _problem(n, p, 21, False)
    g_mode = eng.getLogLikelihoodGradient().reshape(n, p)
    assert np.linalg.norm(g_mode - x) <= 1e-3 * np.linalg.norm(x)  # mode: likelihood + prior gradient = 0
    ceiling = per_location_ceiling(residual_hessian(delta, x, tau) + np.eye(n * p), p)
    assert 0.0 < ceiling <= 1.0
    burn = hmc.default_options(iterations=101, burnIn=100, leapfrogSteps=L, trajectory=0.02 * L, learnPrecision=0,
                               learnTraitPrecision=0, seed=5, adaptStepSize=1)
    _, xb, st = hmc.run_hmc(eng, x, burn)
    sur = Surrogate(n, p, 60 * L + 1, seed=3)
    opt = hmc.default_options(iterations=60, burnIn=0, leapfrogSteps=L, stepSize=st.stepSize, learnPrecision=0,
                              learnTraitPrecision=0, seed=6, surrogate=sur.handle, collectSnapshots=1)
    _, xw, sc = hmc.run_hmc(eng, xb, opt)
    assert sur.snapshot_count() == 1 + L * round(sc.acceptRate * 60)  # start + accepted trajectories
    first = sur.train(1, 2e-3)
    loss = sur.train(8, 2e-3)
    # held-out points: the chain continued, exact gradient rows there
    xs, gs = [], []
    cur = xw
    for k in range(16):
        o = hmc.default_options(iterations=3, burnIn=0, leapfrogSteps=L, stepSize=st.stepSize, learnPrecision=0,
                                learnTraitPrecision=0, seed=100 + k)
        _, cur, _ = hmc.run_hmc(eng, cur, o)
        eng.updateLocations(-1, cur)
        xs.append(cur.copy())
        gs.append(eng.getLogLikelihoodGradient().reshape(n, p))
    gs = np.array(gs)
    xs = np.array(xs)
    pred = np.array([sur.predict(xk) for xk in xs])
    post = gs - xs  # log-posterior gradient rows: the prior part is exact, rotations leave it unchanged
    explained = 1.0 - np.sum((pred - gs) ** 2) / np.sum((post - post.mean(axis=0)) ** 2)
    print(f"tau {tau:.3f} step {st.stepSize:.4g} accept {sc.acceptRate:.2f} train loss {first:.4f} -> {loss:.4f} "
          f"explained variance: trained {explained:.3f}, per-location ceiling {ceiling:.3f}", flush=True)
    assert loss < first
    assert explained > 0.0
    eng.close()
    sur.close()


def test_surrogate_chain_matches_plain_posterior():
    """Surrogate trajectories with the exact accept step (and the fallback)
    leave the posterior unchanged: invariant functionals' means agree with an
    independent plain HMC chain within Monte Carlo error.  The data are
    distances of 2-D points (a well-specified model): with 10-D points the
    2-D posterior has local modes in which two plain chains themselves stay
    apart for thousands of iterations."""
    n, p, L, its = 200, 2, 10, 6000
    delta, eng, x, _ = _problem(n, p, 31, True, dim=2, noise=0.3)
    sur = Surrogate(n, p, 30 * L + L + 1, seed=4)
    opt = hmc.default_options(iterations=31, burnIn=30, leapfrogSteps=L, trajectory=0.02 * L, learnPrecision=0,
                              learnTraitPrecision=0, seed=6, adaptStepSize=1, surrogate=sur.handle,
                              collectSnapshots=1)
    _, xw, st = hmc.run_hmc(eng, x, opt)
    sur.train(6, 2e-3)
    cols = [0] + list(range(3, len(hmc.TRACE_COLUMNS)))
    res = {}
    for name, use, seed in (("plain", 0, 8), ("surrogate", 1, 10)):
        o = hmc.default_options(iterations=its, burnIn=0, leapfrogSteps=L, stepSize=st.stepSize, learnPrecision=0,
                                learnTraitPrecision=0, seed=seed, surrogate=sur.handle, useSurrogate=use,
                                fallbackAcceptRate=0.2, fallbackWindow=10)
        trace, _, s = hmc.run_hmc(eng, xw, o)
        t = trace[its // 10:, cols]
        ess = np.array([hmc.effective_sample_size(t[:, k]) for k in range(t.shape[1])])
        res[name] = (t.mean(axis=0), t.std(axis=0) / np.sqrt(np.maximum(ess, 1.0)), s)
        print(f"{name}: accept {s.acceptRate:.3f} surrogate its {s.surrogateIterations:.0f} "
              f"exact its {s.exactIterations:.0f} seconds {s.samplingSeconds:.3f}", flush=True)
    (m0, se0, _), (m1, se1, s1) = res["plain"], res["surrogate"]
    z = np.abs(m0 - m1) / np.hypot(se0, se1)
    print("z", np.round(z, 2), flush=True)
    assert s1.surrogateIterations > 0
    assert s1.surrogateIterations + s1.exactIterations == its
    assert np.all(z < 3.5)
    eng.close()
    sur.close()


def _log_posterior(eng, x, precision, truncation):
    """Exact log posterior of the sampler's model at x (trait precision I,
    row covariance I_N): the engine's log-likelihood (R/mds.R:24-28) plus the
    matrix-normal log prior (R/Untitled.R:68-95)."""
    n, p = x.shape
    eng.updateLocations(-1, x)
    s = eng.getSumOfIncrements()
    loglik = 0.5 * (np.log(precision) - np.log(2.0 * np.pi)) * n * (n - 1) / 2.0 - s
    return loglik - 0.5 * float(np.sum(x * x)) - 0.5 * n * p * np.log(2.0 * np.pi)


@pytest.mark.timeout(900)
def test_config5_surrogate_sampler_at_full_size():
    """The surrogate-trajectory sampler at its workload's shape: N = 20,000,
    P = 5, distances of N(0, I_10) points with 2% missing, truncation, the
    5 + 16 -> 256 -> 256 -> 5 network, 2,000 warm-up evaluations.  Checked
    through properties that hold at any size:
      * the warm-up stores 1 + L snapshots per accepted trajectory (a
        rejected trajectory's snapshots are dropped);
      * the fallback schedule: with every window failing (threshold above 1)
        windows of W surrogate iterations alternate with exact windows of
        W, 2W, 4W iterations;
      * the accept step judges with the exact likelihood: after runs that
        accept surrogate trajectories, the log posterior the trace records
        for the returned state equals a fresh exact evaluation there;
      * every trace entry is finite."""
    n, p, L = 20000, 5, 20
    print(f"config 5: N={n} P={p}: data", flush=True)
    delta = datagen.euclid10(n, missing=0.02)
    eng = MultiDimensionalScaling(p, n, Flags.OPENCL | Flags.LEFT_TRUNCATION, 0)
    eng.setPairwiseData(delta)
    eng.setParameters([1.0])
    x = map_locations(eng, datagen.normals(666, 3, n * p).reshape(n, p), steps=60)
    tau = plugin_precision(delta, x)
    del delta
    eng.setParameters([tau])
    print(f"config 5: precision {tau:.4f}; warm-up", flush=True)
    warm_evals = 2000
    sur = Surrogate(n, p, warm_evals + L + 1, seed=5)
    its = warm_evals // L
    opt = hmc.default_options(iterations=its + 1, burnIn=its, leapfrogSteps=L, trajectory=0.01 * L,
                              learnPrecision=0, learnTraitPrecision=0, seed=7, adaptStepSize=1, targetAccept=0.8,
                              surrogate=sur.handle, collectSnapshots=1)
    trace, xw, st = hmc.run_hmc(eng, x, opt)
    accepted = round(st.acceptRate * (its + 1))
    assert st.snapshots == 1 + L * accepted and sur.snapshot_count() == st.snapshots
    assert accepted > 0 and np.all(np.isfinite(trace))
    loss = sur.train(4, 2e-3)
    assert np.isfinite(loss)
    eps = st.stepSize
    # fallback schedule: every window fails
    w = 5
    o = hmc.default_options(iterations=10 * w, burnIn=0, leapfrogSteps=L, stepSize=eps, learnPrecision=0,
                            learnTraitPrecision=0, seed=8, surrogate=sur.handle, useSurrogate=1,
                            fallbackAcceptRate=1.01, fallbackWindow=w)
    trace, xf, s = hmc.run_hmc(eng, xw, o)
    print(f"config 5: fallback schedule surrogate {s.surrogateIterations:.0f} exact {s.exactIterations:.0f}",
          flush=True)
    assert (s.surrogateIterations, s.exactIterations) == (3 * w, 7 * w)
    assert np.all(np.isfinite(trace))
    # accepted surrogate trajectories (short, small steps), exact energies of the returned states
    cur, surrogate_accepts = xf, 0
    for k in range(6):
        o = hmc.default_options(iterations=2, burnIn=0, leapfrogSteps=2, stepSize=0.05 * eps, learnPrecision=0,
                                learnTraitPrecision=0, seed=20 + k, surrogate=sur.handle, useSurrogate=1,
                                fallbackWindow=0)
        trace, cur, s = hmc.run_hmc(eng, cur, o)
        assert np.all(np.isfinite(trace))
        surrogate_accepts += round(s.surrogateAcceptRate * s.surrogateIterations)
        exact = _log_posterior(eng, cur, tau, True)
        assert abs(trace[-1, 0] - exact) <= 1e-12 * abs(exact), (k, trace[-1, 0], exact)
    print(f"config 5: accepted surrogate trajectories {surrogate_accepts} of 12", flush=True)
    assert surrogate_accepts > 0
    eng.close()
    sur.close()


@pytest.mark.timeout(900)
def test_surrogate_chain_matches_plain_posterior_p5():
    """Exactness at P = 5: on a posterior where two independent plain chains
    agree (distances of 5-D points, N = 120, so the model is well specified),
    the surrogate chain's invariant functionals agree with them within
    Monte Carlo error."""
    n, p, L, its = 120, 5, 10, 6000
    delta, eng, x, _ = _problem(n, p, 41, True, dim=5, noise=0.3)
    sur = Surrogate(n, p, 30 * L + L + 1, seed=4)
    opt = hmc.default_options(iterations=31, burnIn=30, leapfrogSteps=L, trajectory=0.02 * L, learnPrecision=0,
                              learnTraitPrecision=0, seed=6, adaptStepSize=1, surrogate=sur.handle,
                              collectSnapshots=1)
    _, xw, st = hmc.run_hmc(eng, x, opt)
    sur.train(6, 2e-3)
    cols = [0] + list(range(3, len(hmc.TRACE_COLUMNS)))
    res = {}
    for name, use, seed in (("plain", 0, 8), ("plain2", 0, 9), ("surrogate", 1, 10)):
        o = hmc.default_options(iterations=its, burnIn=0, leapfrogSteps=L, stepSize=st.stepSize, learnPrecision=0,
                                learnTraitPrecision=0, seed=seed, surrogate=sur.handle, useSurrogate=use,
                                fallbackAcceptRate=0.2, fallbackWindow=10)
        trace, _, s = hmc.run_hmc(eng, xw, o)
        t = trace[its // 10:, cols]
        ess = np.array([hmc.effective_sample_size(t[:, k]) for k in range(t.shape[1])])
        res[name] = (t.mean(axis=0), t.std(axis=0) / np.sqrt(np.maximum(ess, 1.0)), s)
        print(f"{name}: accept {s.acceptRate:.3f} surrogate its {s.surrogateIterations:.0f} min ESS "
              f"{np.nanmin(ess):.0f}", flush=True)
    m0, se0, _ = res["plain"]
    m2, se2, _ = res["plain2"]
    m1, se1, s1 = res["surrogate"]
    z_plain = np.abs(m0 - m2) / np.hypot(se0, se2)
    z = np.abs(m0 - m1) / np.hypot(se0, se1)
    print("z plain-plain", np.round(z_plain, 2), "z plain-surrogate", np.round(z, 2), flush=True)
    assert np.all(z_plain < 3.5)  # the two plain chains agree: the comparison is meaningful
    assert s1.surrogateIterations > 0
    assert np.all(z < 3.5)
    eng.close()
    sur.close()
