"""Per-location MLP surrogate of the log-likelihood gradient rows (C ABI:
mds_hip_surrogate_* in include/mds_hip.h) and the surrogate-trajectory HMC
workflow built on it:

1. plain-gradient HMC warm-up with ``collectSnapshots``: every exact gradient
   pass stores (x, gradient) on the device;
2. ``train``: the MLP (BF16 weights on MFMA, FP32 accumulation, Adam) fits the
   stored gradient rows;
3. HMC with ``useSurrogate``: leapfrog steps use the MLP's gradients, the
   exact FP64-reduced likelihood decides acceptance, and exact gradients take
   over for a window whenever the acceptance over a window drops below a
   threshold.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .engine import MdsError, _c_double_p, load_library

HIDDEN = 256
EMBED = 16
IN = 32
OUT = 16

_bound = False


def _bind(lib):
    global _bound
    if _bound:
        return lib
    H = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "mds_hip_surrogate_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "mds_hip_surrogate_destroy": ([H], ctypes.c_int),
        "mds_hip_surrogate_last_error": ([H], ctypes.c_char_p),
        "mds_hip_surrogate_snapshot_count": ([H, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "mds_hip_surrogate_add_snapshot": ([H, _c_double_p, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_surrogate_train": ([H, ctypes.c_int, ctypes.c_double, _c_double_p], ctypes.c_int),
        "mds_hip_surrogate_predict": ([H, _c_double_p, ctypes.c_size_t, _c_double_p], ctypes.c_int),
        "mds_hip_surrogate_parameter_count": ([H, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "mds_hip_surrogate_get_parameters": ([H, fp, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_surrogate_set_parameters": ([H, fp, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_surrogate_set_normalisation": ([H, fp, fp, ctypes.c_size_t, fp, fp, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_surrogate_loss_gradient": ([H, _c_double_p, _c_double_p, ctypes.c_size_t, fp, ctypes.c_size_t,
                                             _c_double_p], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _bound = True
    return lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Surrogate:
    """MLP surrogate for N locations in P dimensions with room for
    ``capacity`` training snapshots."""

    def __init__(self, locationCount: int, embeddingDimension: int, capacity: int, seed: int = 1, device: int = 0):
        self._lib = _bind(load_library())
        h = ctypes.c_void_p()
        st = self._lib.mds_hip_surrogate_create(int(locationCount), int(embeddingDimension), int(capacity),
                                                int(seed), int(device), ctypes.byref(h))
        if st != 0:
            raise MdsError(st, self._lib.mds_hip_surrogate_last_error(None).decode())
        self._h = h
        self.n = int(locationCount)
        self.p = int(embeddingDimension)

    @property
    def handle(self):
        return self._h

    def _check(self, st):
        if st != 0:
            raise MdsError(st, self._lib.mds_hip_surrogate_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mds_hip_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def snapshot_count(self) -> int:
        c = ctypes.c_int()
        self._check(self._lib.mds_hip_surrogate_snapshot_count(self._h, ctypes.byref(c)))
        return c.value

    def add_snapshot(self, x, g) -> None:
        x, g = _f64(x).ravel(), _f64(g).ravel()
        self._check(self._lib.mds_hip_surrogate_add_snapshot(self._h, x.ctypes.data_as(_c_double_p),
                                                             g.ctypes.data_as(_c_double_p), x.size))

    def train(self, epochs: int, learning_rate: float = 1e-3) -> float:
        loss = ctypes.c_double()
        self._check(self._lib.mds_hip_surrogate_train(self._h, int(epochs), float(learning_rate),
                                                      ctypes.byref(loss)))
        return loss.value

    def predict(self, x) -> np.ndarray:
        x = _f64(x).ravel()
        out = np.empty_like(x)
        self._check(self._lib.mds_hip_surrogate_predict(self._h, x.ctypes.data_as(_c_double_p), x.size,
                                                        out.ctypes.data_as(_c_double_p)))
        return out.reshape(self.n, self.p)

    def parameter_count(self) -> int:
        c = ctypes.c_longlong()
        self._check(self._lib.mds_hip_surrogate_parameter_count(self._h, ctypes.byref(c)))
        return c.value

    def get_parameters(self) -> np.ndarray:
        out = np.empty(self.parameter_count(), dtype=np.float32)
        self._check(self._lib.mds_hip_surrogate_get_parameters(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size))
        return out

    def set_parameters(self, params) -> None:
        a = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self._lib.mds_hip_surrogate_set_parameters(
            self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size))

    def set_normalisation(self, x_centre, g_centre, sigma_x, sigma_g) -> None:
        """Per-location centres (N x P) of inputs and targets, per-dimension
        scales (P)."""
        fp = ctypes.POINTER(ctypes.c_float)
        xc, gc, sx, sg = [np.ascontiguousarray(v, dtype=np.float32).ravel()
                          for v in (x_centre, g_centre, sigma_x, sigma_g)]
        self._check(self._lib.mds_hip_surrogate_set_normalisation(
            self._h, xc.ctypes.data_as(fp), gc.ctypes.data_as(fp), xc.size, sx.ctypes.data_as(fp),
            sg.ctypes.data_as(fp), sx.size))

    def loss_gradient(self, x, g):
        """(loss, parameter gradient) of one forward + backward, no update."""
        x, g = _f64(x).ravel(), _f64(g).ravel()
        grad = np.empty(self.parameter_count(), dtype=np.float32)
        loss = ctypes.c_double()
        self._check(self._lib.mds_hip_surrogate_loss_gradient(
            self._h, x.ctypes.data_as(_c_double_p), g.ctypes.data_as(_c_double_p), x.size,
            grad.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), grad.size, ctypes.byref(loss)))
        return loss.value, grad


def split_parameters(params, n):
    """Views of the flat parameter vector: W1, b1, W2, b2, W3, b3, E."""
    shapes = [(IN, HIDDEN), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (HIDDEN, OUT), (OUT,), (n, EMBED)]
    out, k = [], 0
    for s in shapes:
        size = int(np.prod(s))
        out.append(params[k:k + size].reshape(s))
        k += size
    return out


def residual_hessian(delta, x, precision: float) -> np.ndarray:
    """Minus the Jacobian (NP x NP) of the untruncated gradient rows the
    engine returns, g_i = sum over observed delta_ij of tau r_ij u_ij
    (src/NewMultiDimensionalScaling.hpp:390-446 row semantics, r = delta - d,
    u the unit vector from x_j to x_i): the pair block
    A_ij = tau (u u^T - (r / d) (I - u u^T)) enters H_ii with + and H_ij
    with -.  Symmetric, and the Hessian of minus the log-likelihood, when the
    observations are symmetric.  Dense: for diagnostics at small N."""
    delta = np.asarray(delta, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    n, p = x.shape
    obs = ~np.isnan(delta)
    np.fill_diagonal(obs, False)
    diff = x[:, None, :] - x[None, :, :]
    d = np.sqrt((diff ** 2).sum(-1))
    np.fill_diagonal(d, 1.0)
    u = diff / d[:, :, None]
    r = np.where(obs, delta - d, 0.0)
    uu = u[:, :, :, None] * u[:, :, None, :]
    a = precision * obs[:, :, None, None] * (uu - (r / d)[:, :, None, None] * (np.eye(p) - uu))
    h = -a.transpose(0, 2, 1, 3).reshape(n * p, n * p)
    diag = a.sum(axis=1)
    for i in range(n):
        h[i * p:(i + 1) * p, i * p:(i + 1) * p] = diag[i]
    return h


def per_location_ceiling(hessian, p: int, flat: float = 1e-8) -> float:
    """Under a Gaussian approximation of the posterior with precision H
    (NP x NP), the fraction of the posterior variance of the log-posterior
    gradient that the best model of row i from x_i alone can explain.  With
    Sigma the pseudo-inverse of H and Pi the projection onto H's range
    (eigenvalues above ``flat`` times the largest: the flat directions, such
    as the rotations of a rotation-invariant posterior, are left out),
    Cov(grad) = H and Cov(grad, x) = -Pi, so the regression of row i on x_i
    explains Pi_ii Sigma_ii^{-1} Pi_ii out of H_ii, and the fraction is the
    ratio of the traces summed over i.  It is 1 when H is block diagonal and
    drops with the coupling between locations.  A non-symmetric H
    (asymmetric observations) enters by its symmetric part."""
    h = np.asarray(hessian, dtype=np.float64)
    h = 0.5 * (h + h.T)
    n = h.shape[0] // p
    w, v = np.linalg.eigh(h)
    keep = w > flat * w.max()
    vk = v[:, keep]
    sigma = (vk / w[keep]) @ vk.T
    proj = vk @ vk.T
    num = 0.0
    for i in range(n):
        blk = slice(i * p, (i + 1) * p)
        pi = proj[blk, blk]
        num += np.trace(pi @ np.linalg.pinv(sigma[blk, blk]) @ pi)
    den = sum(np.trace(h[i * p:(i + 1) * p, i * p:(i + 1) * p]) for i in range(n))
    return float(num / den)


def map_locations(engine, x0, steps: int = 500, trait=None, tolerance: float = 1e-10):
    """Posterior mode: L-BFGS (scipy) on minus the log posterior up to a
    constant, S(x) + tr(X T X^T) / 2 with S the engine's sum of increments
    and T the trait precision (default I; row covariance I_N), with the
    engine's exact gradient.  A warm start shared by the samplers being
    compared."""
    from scipy.optimize import minimize

    x0 = np.array(x0, dtype=np.float64)
    n, p = x0.shape
    t = np.eye(p) if trait is None else np.asarray(trait, dtype=np.float64)

    def objective(flat):
        x = flat.reshape(n, p)
        engine.updateLocations(-1, x)
        s = engine.getSumOfIncrements()
        g = engine.getLogLikelihoodGradient().reshape(n, p)
        xt = x @ t
        return s + 0.5 * float(np.sum(x * xt)), (xt - g).ravel()

    res = minimize(objective, x0.ravel(), jac=True, method="L-BFGS-B",
                   options={"maxiter": steps, "maxcor": 20, "ftol": tolerance, "gtol": 1e-9})
    x = res.x.reshape(n, p)
    engine.updateLocations(-1, x)
    return x


def plugin_precision(delta, x) -> float:
    """Residual precision at locations x: observed pairs over the sum of
    squared residuals (delta_ij - d_ij)^2, from an untruncated engine at
    precision 1, whose sum of increments is half that sum for symmetric
    observations."""
    from .engine import Flags, MultiDimensionalScaling

    delta = np.asarray(delta, dtype=np.float64)
    n, p = np.asarray(x).shape
    obs = ~np.isnan(delta)
    np.fill_diagonal(obs, False)
    with MultiDimensionalScaling(p, n, Flags.OPENCL, 0) as e:
        e.setPairwiseData(delta)
        e.setParameters([1.0])
        e.updateLocations(-1, x)
        s = e.getSumOfIncrements()
    return float(0.5 * obs.sum() / (2.0 * s))


def _chain_summary(trace, st, iterations):
    """Accept rate, time, min-ESS and min-ESS/s of a fixed-precision run over
    the invariant functionals (log posterior, distances, norms; a chain that
    never moved has ESS 0)."""
    from . import hmc

    cols = [0] + list(range(3, len(hmc.TRACE_COLUMNS)))  # tau and trace(T) are fixed here
    ess = np.array([hmc.effective_sample_size(trace[:, k]) for k in cols])
    ess = np.where(np.isfinite(ess), ess, 0.0)
    means = trace[:, cols].mean(axis=0)
    se = trace[:, cols].std(axis=0) / np.sqrt(np.maximum(ess, 1.0))
    return {
        "accept_rate": st.acceptRate, "seconds": st.samplingSeconds,
        "ms_per_iteration": st.samplingSeconds / iterations * 1e3, "min_ess": float(ess.min()),
        "min_ess_per_s": float(ess.min()) / st.samplingSeconds if st.samplingSeconds > 0 else 0.0,
        "surrogate_iterations": st.surrogateIterations, "exact_iterations": st.exactIterations,
        "surrogate_accept_rate": st.surrogateAcceptRate, "step_size": st.stepSize,
        "means": means, "se": se, "ess": ess,
    }


def sweep_surrogate(eng, sur, x_start, precision, step0: float, leapfrogs=(2, 5, 10, 20), burn_in: int = 200,
                    iterations: int = 2000, target_accept: float = 0.65, seed: int = 900):
    """The surrogate chain tuned on its own: for each leapfrog count L, a
    burn-in of surrogate trajectories dual-averages the chain's own step size,
    from ``step0``, toward ``target_accept`` (no exact-gradient fallback, so only surrogate
    trajectories are proposed), then ``iterations`` more are sampled at the
    adapted step.  Returns one summary per L (acceptance, step size, time per
    iteration, min-ESS and min-ESS/s over the invariant functionals)."""
    from . import hmc

    out = []
    for k, L in enumerate(leapfrogs):
        eng.setParameters([precision])
        o = hmc.default_options(iterations=burn_in + iterations, burnIn=burn_in, leapfrogSteps=int(L),
                                trajectory=step0 * L, learnPrecision=0, learnTraitPrecision=0, seed=seed + k,
                                adaptStepSize=1, targetAccept=target_accept, surrogate=sur.handle, useSurrogate=1,
                                fallbackWindow=0)
        trace, _, st = hmc.run_hmc(eng, x_start, o)
        r = _chain_summary(trace, st, iterations)
        r["leapfrog"] = int(L)
        out.append(r)
    return out


def grid_surrogate(eng, sur, x_start, precision, step0: float, leapfrogs=(2, 5, 10, 20),
                   fractions=(1.0, 0.25, 0.0625, 0.015625), iterations: int = 400, seed: int = 950):
    """Acceptance of surrogate trajectories over a grid of fixed step sizes
    (``fractions`` of ``step0``) and leapfrog counts, with no exact-gradient
    fallback: where the gradient error, not the integrator, sets the
    acceptance, the acceptance follows the trajectory length (step x L)
    whatever the split.  The runs are short: their ESS figures are not
    meaningful, only their acceptance."""
    from . import hmc

    out = []
    k = 0
    for L in leapfrogs:
        for f in fractions:
            eng.setParameters([precision])
            o = hmc.default_options(iterations=iterations, burnIn=0, leapfrogSteps=int(L), stepSize=step0 * f,
                                    learnPrecision=0, learnTraitPrecision=0, seed=seed + k, surrogate=sur.handle,
                                    useSurrogate=1, fallbackWindow=0)
            trace, _, st = hmc.run_hmc(eng, x_start, o)
            r = _chain_summary(trace, st, iterations)
            r["leapfrog"] = int(L)
            r["trajectory"] = step0 * f * L
            out.append(r)
            k += 1
    return out


def compare_samplers(delta, p: int, iterations: int, warmup_evals: int = 2000, leapfrog: int = 20,
                     truncation: bool = True, precision=None, epochs: int = 4, learning_rate: float = 2e-3,
                     fallback_rate: float = 0.2, fallback_window: int = 20, seed: int = 666, map_steps: int = 300,
                     sweep=None, sweep_iterations: int = 2000, grid=False, log=None):
    """Plain-gradient HMC against surrogate-trajectory HMC on the same target
    (fixed precision and trait precision), from the same warm start: the
    plain chain at the step size dual-averaged during the warm-up, the
    surrogate chain first at the same step and leapfrog count (with the
    exact-gradient fallback), then, with ``sweep`` (leapfrog counts), tuned on
    its own by ``sweep_surrogate``.  Returns a dict with the runs' statistics,
    the min-ESS/s ratio of the fallback run and of the best tuned run, and
    the largest difference of posterior means in Monte Carlo standard
    errors.  ``log``, when given, receives a line at the start of each
    stage."""
    from . import datagen, hmc

    say = log or (lambda message: None)
    from .engine import Flags, MultiDimensionalScaling

    delta = np.asarray(delta, dtype=np.float64)
    n = delta.shape[0]
    flags = Flags.OPENCL | (Flags.LEFT_TRUNCATION if truncation else 0)
    eng = MultiDimensionalScaling(p, n, flags, 0)
    eng.setPairwiseData(delta)
    eng.setParameters([1.0 if precision is None else float(precision)])
    import time

    t0 = time.perf_counter()
    say("posterior mode")
    x = map_locations(eng, datagen.normals(seed, 3, n * p).reshape(n, p), steps=map_steps)
    if precision is None:
        precision = plugin_precision(delta, x)
        eng.setParameters([precision])
        x = map_locations(eng, x, steps=map_steps)
    t_map = time.perf_counter() - t0

    sur = Surrogate(n, p, warmup_evals + leapfrog + 1, seed=seed)
    warm_its = max(1, warmup_evals // leapfrog)
    opt = hmc.default_options(iterations=warm_its + 1, burnIn=warm_its, leapfrogSteps=leapfrog,
                              trajectory=0.01 * leapfrog, learnPrecision=0, learnTraitPrecision=0, seed=seed,
                              adaptStepSize=1, targetAccept=0.8, surrogate=sur.handle, collectSnapshots=1)
    t0 = time.perf_counter()
    say("warm-up with snapshots")
    _, x_warm, st_warm = hmc.run_hmc(eng, x, opt)
    t_warm = time.perf_counter() - t0
    eps = st_warm.stepSize

    t0 = time.perf_counter()
    say("training")
    loss = sur.train(epochs, learning_rate)
    t_train = time.perf_counter() - t0

    runs = {}
    for name, use, chain_seed in (("plain", 0, seed + 1), ("surrogate", 1, seed + 2)):  # independent chains
        eng.setParameters([precision])
        say(f"{name} chain")
        o = hmc.default_options(iterations=iterations, burnIn=0, leapfrogSteps=leapfrog, stepSize=eps,
                                learnPrecision=0, learnTraitPrecision=0, seed=chain_seed, surrogate=sur.handle,
                                useSurrogate=use, fallbackAcceptRate=fallback_rate,
                                fallbackWindow=fallback_window)
        trace, _, st = hmc.run_hmc(eng, x_warm, o)
        runs[name] = _chain_summary(trace, st, iterations)

    def zmax(a, b):
        z = np.abs(a["means"] - b["means"]) / np.hypot(a["se"], b["se"])
        return float(np.max(z))

    say("tuned surrogate chains")
    tuned = sweep_surrogate(eng, sur, x_warm, precision, eps, sweep, iterations=sweep_iterations) if sweep else []
    for r in tuned:
        r["max_mean_z"] = zmax(runs["plain"], r)
    # a rate counts only where the chain's means match the plain chain's
    # within 2 Monte Carlo standard errors: a chain that has not mixed has
    # no meaningful ESS
    matched = [r for r in tuned if r["max_mean_z"] <= 2.0]
    best = max(matched, key=lambda r: r["min_ess_per_s"]) if matched else None
    if grid:
        say("fixed-step grid")
    grid_runs = grid_surrogate(eng, sur, x_warm, precision, eps) if grid else []
    eng.close()
    sur.close()
    plain_rate = runs["plain"]["min_ess_per_s"]
    return {
        "n": n, "p": p, "precision": precision, "step_size": eps, "warmup_accept": st_warm.acceptRate,
        "snapshots": st_warm.snapshots, "train_loss": loss, "map_s": t_map, "warmup_s": t_warm,
        "train_s": t_train, "plain": runs["plain"], "surrogate": runs["surrogate"],
        "ess_per_s_ratio": runs["surrogate"]["min_ess_per_s"] / plain_rate,
        "max_mean_z": zmax(runs["plain"], runs["surrogate"]),
        "tuned": tuned,
        "grid": grid_runs,
        "best_tuned_ratio": best["min_ess_per_s"] / plain_rate if best else None,
        "best_tuned_z": best["max_mean_z"] if best else None,
        "columns": [hmc.TRACE_COLUMNS[k] for k in [0] + list(range(3, len(hmc.TRACE_COLUMNS)))],
    }
