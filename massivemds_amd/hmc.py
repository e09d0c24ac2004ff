"""HMC for Bayesian MDS on the GPU: the Python counterpart of the reference's
R sampler ``hmcsampler`` (R/Untitled.R:477-729), driving the device-resident
sampler in libmds_hip.so (``mds_hip_hmc_run``), plus effective-sample-size
estimation on rotation/reflection-invariant functionals.

Differences from the R function, by design: no BEAST tree input (the row
covariance is I_N, the R path taken when ``beast`` is NULL); the
precision's update is Metropolis under truncation and a Gamma Gibbs draw
without it; the potential is recomputed after each precision or trait
precision update; the precision step uses the current (accepted) locations.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import datagen
from .engine import Flags, MdsError, MultiDimensionalScaling, _c_double_p, load_library

TRACE_COLUMNS = ["log_posterior", "precision", "trace_trait_precision", "d_0_1", "d_0_half", "d_third_2third",
                 "d_1_last", "d_quarter_3quarter", "norm_0", "norm_half", "norm_last"]


class HmcOptions(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("burnIn", ctypes.c_int),
        ("leapfrogSteps", ctypes.c_int),
        ("trajectory", ctypes.c_double),
        ("learnPrecision", ctypes.c_int),
        ("learnTraitPrecision", ctypes.c_int),
        ("traitInvWeight", ctypes.c_int),
        ("precisionProposalWidth", ctypes.c_double),
        ("priorDf", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
        ("useGraph", ctypes.c_int),
        ("adaptStepSize", ctypes.c_int),
        ("targetAccept", ctypes.c_double),
        ("stepSize", ctypes.c_double),
        ("surrogate", ctypes.c_void_p),
        ("useSurrogate", ctypes.c_int),
        ("collectSnapshots", ctypes.c_int),
        ("fallbackAcceptRate", ctypes.c_double),
        ("fallbackWindow", ctypes.c_int),
        ("adaptPrecisionWidth", ctypes.c_int),
    ]


class HmcStats(ctypes.Structure):
    _fields_ = [
        ("acceptRate", ctypes.c_double),
        ("precisionAcceptRate", ctypes.c_double),
        ("samplingSeconds", ctypes.c_double),
        ("totalSeconds", ctypes.c_double),
        ("evaluations", ctypes.c_double),
        ("finalPrecision", ctypes.c_double),
        ("stepSize", ctypes.c_double),
        ("finalPotential", ctypes.c_double),
        ("surrogateIterations", ctypes.c_double),
        ("exactIterations", ctypes.c_double),
        ("snapshots", ctypes.c_double),
        ("surrogateAcceptRate", ctypes.c_double),
        ("precisionProposalWidth", ctypes.c_double),
    ]


def _bind(lib):
    lib.mds_hip_hmc_default_options.argtypes = [ctypes.POINTER(HmcOptions)]
    lib.mds_hip_hmc_default_options.restype = ctypes.c_int
    lib.mds_hip_hmc_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(HmcOptions), _c_double_p, ctypes.c_size_t,
                                    _c_double_p, ctypes.c_size_t, _c_double_p, ctypes.POINTER(HmcStats)]
    lib.mds_hip_hmc_run.restype = ctypes.c_int
    return lib


def default_options(**overrides) -> HmcOptions:
    opt = HmcOptions()
    _bind(load_library()).mds_hip_hmc_default_options(ctypes.byref(opt))
    for k, v in overrides.items():
        setattr(opt, k, v)
    return opt


def run_hmc(engine: MultiDimensionalScaling, initial, options: HmcOptions):
    """Runs the device sampler on an engine holding data and a precision."""
    lib = _bind(load_library())
    x0 = np.ascontiguousarray(initial, dtype=np.float64).ravel()
    kept = options.iterations - options.burnIn
    trace = np.empty(kept * len(TRACE_COLUMNS))
    final = np.empty(engine.locationCount * engine.embeddingDimension)
    stats = HmcStats()
    st = lib.mds_hip_hmc_run(engine._h, ctypes.byref(options), x0.ctypes.data_as(_c_double_p), x0.size,
                             trace.ctypes.data_as(_c_double_p), trace.size, final.ctypes.data_as(_c_double_p),
                             ctypes.byref(stats))
    if st != 0:
        raise MdsError(st, lib.mds_hip_last_error(engine._h).decode())
    return (trace.reshape(kept, len(TRACE_COLUMNS)),
            final.reshape(engine.locationCount, engine.embeddingDimension), stats)


def effective_sample_size(x) -> float:
    """Geyer's initial monotone sequence estimator (Geyer 1992, "Practical
    Markov chain Monte Carlo", Statist. Sci. 7): autocorrelations by FFT,
    sums of adjacent pairs truncated at the first non-positive pair and made
    monotone; ESS = n / (-1 + 2 sum of the pair sums)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    if n < 4 or np.ptp(x) == 0.0:
        return float("nan")
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    acov = np.fft.irfft(f * np.conj(f))[:n] / n
    rho = acov / acov[0]
    pairs = rho[0:n - 1:2][: (n - 1) // 2] + rho[1:n:2][: (n - 1) // 2]
    total = 0.0
    prev = math.inf
    for g in pairs:
        if g <= 0.0:
            break
        g = min(g, prev)
        total += g
        prev = g
    tau = -1.0 + 2.0 * total
    return float(n / max(tau, 1e-12))


def hmcsampler(n_iter: int, burnIn: int = 0, data=None, latentDimension: int = 2, learnPrec: bool = False,
               learnTraitPrec: bool = False, trajectory: float = 0.2, traitInvWeight: int = 1,
               mdsPrecision: float = 1.0, truncation: bool = True, gpu: int = 1, single: bool = False,
               leapfrog: int = 20, seed: int = 666, useGraph: bool = True, precisionWidth: float = 0.01,
               initial=None, adaptPrecisionWidth: bool = False):
    """R/Untitled.R:477-729 on the GPU.  Returns a dict with the trace of the
    invariant functionals after burn-in (``target`` = minus the log posterior,
    as R's Target), ``Time`` (seconds after burn-in), ``acprat`` (acceptance
    rate), ``precision``, ``locations`` (final), ``ess`` per functional and
    ``min_ess_per_s``.  ``adaptPrecisionWidth`` adapts the width of the
    precision's random-walk proposal during burn-in (the reference keeps
    ``precisionWidth`` fixed)."""
    if data is None:
        raise ValueError("No data found.")
    if gpu <= 0:
        raise ValueError("the sampler runs on the GPU engine: gpu >= 1")
    data = np.asarray(data, dtype=np.float64)
    n = data.shape[0]
    p = int(latentDimension)
    flags = Flags.OPENCL | (Flags.LEFT_TRUNCATION if truncation else 0) | (Flags.FLOAT if single else 0)
    eng = MultiDimensionalScaling(p, n, flags, gpu - 1 if gpu > 1 else 0)
    eng.setPairwiseData(data)
    eng.setParameters([mdsPrecision])
    x0 = datagen.normals(seed, 3, n * p).reshape(n, p) if initial is None else np.asarray(initial)
    opt = HmcOptions()
    _bind(load_library()).mds_hip_hmc_default_options(ctypes.byref(opt))
    opt.iterations = int(n_iter)
    opt.burnIn = int(burnIn)
    opt.leapfrogSteps = int(leapfrog)
    opt.trajectory = float(trajectory)
    opt.learnPrecision = int(bool(learnPrec))
    opt.learnTraitPrecision = int(bool(learnTraitPrec))
    opt.traitInvWeight = int(traitInvWeight)
    opt.precisionProposalWidth = float(precisionWidth)
    opt.adaptPrecisionWidth = int(bool(adaptPrecisionWidth))
    opt.seed = int(seed)
    opt.useGraph = int(bool(useGraph))
    trace, final, stats = run_hmc(eng, x0, opt)
    eng.close()
    ess = {c: effective_sample_size(trace[:, k]) for k, c in enumerate(TRACE_COLUMNS)}
    finite = [v for c, v in ess.items() if np.isfinite(v) and not (c == "precision" and not learnPrec)
              and not (c == "trace_trait_precision" and not learnTraitPrec)]
    min_ess = min(finite) if finite else float("nan")
    return {
        "trace": trace, "columns": TRACE_COLUMNS, "target": -trace[:, 0], "Time": stats.samplingSeconds,
        "acprat": stats.acceptRate, "precision_acceptance": stats.precisionAcceptRate,
        "precision": trace[:, 1], "locations": final, "ess": ess, "min_ess": min_ess,
        "min_ess_per_s": min_ess / stats.samplingSeconds if stats.samplingSeconds > 0 else float("nan"),
        "evaluations": stats.evaluations, "total_seconds": stats.totalSeconds, "step_size": stats.stepSize,
        "precision_width": stats.precisionProposalWidth,
    }
