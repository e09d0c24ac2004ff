"""Python access to the oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker; the product (massivemds_amd/) never
imports it.

* ``liboracle.so`` (built by oracle/Makefile from mds_oracle.c and
  mds_oracle_simd.c) is the C restatement of the reference serial engine
  (src/NewMultiDimensionalScaling.hpp:134-561) and the multi-core SIMD CPU
  baseline.
* ``closed_form_loglik_mp`` restates the reference's R closed form
  ``computeLoglikelihood`` (R/Untitled.R:16-52) in mpmath at high precision,
  extended with the engine's missing-pair rule; it is an independent
  formulation (normal log density and log CDF summed over i < j) used to
  pin the engine arithmetic, and its finite differences pin the gradient.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = ctypes.CDLL(str(LIB))
    i64 = ctypes.c_int64
    lib.mds_oracle_sum_of_increments_f64.argtypes = [_dp, _dp, i64, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    lib.mds_oracle_sum_of_increments_f64.restype = ctypes.c_double
    lib.mds_oracle_sum_of_increments_f32.argtypes = lib.mds_oracle_sum_of_increments_f64.argtypes
    lib.mds_oracle_sum_of_increments_f32.restype = ctypes.c_double
    lib.mds_oracle_gradient_f64.argtypes = [_dp, _dp, i64, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
    lib.mds_oracle_gradient_f64.restype = None
    lib.mds_oracle_gradient_f32.argtypes = lib.mds_oracle_gradient_f64.argtypes
    lib.mds_oracle_gradient_f32.restype = None
    lib.mds_oracle_rows_sum_of_increments_f64.argtypes = [_dp, _dp, i64, ctypes.c_int, ctypes.c_double,
                                                          ctypes.c_int, i64]
    lib.mds_oracle_rows_sum_of_increments_f64.restype = ctypes.c_double
    lib.mds_oracle_rows_gradient_f64.argtypes = [_dp, _dp, i64, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                 i64, _dp]
    lib.mds_oracle_rows_gradient_f64.restype = None
    lib.mds_oracle_log_likelihood.argtypes = [ctypes.c_double, i64, ctypes.c_double]
    lib.mds_oracle_log_likelihood.restype = ctypes.c_double
    lib.mds_oracle_simd_sum_of_increments.argtypes = [_dp, _dp, i64, ctypes.c_int, ctypes.c_double,
                                                      ctypes.c_int, ctypes.c_int, i64]
    lib.mds_oracle_simd_sum_of_increments.restype = ctypes.c_double
    lib.mds_oracle_simd_gradient.argtypes = [_dp, _dp, i64, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_int, i64, _dp]
    lib.mds_oracle_simd_gradient.restype = None
    lib.mds_oracle_simd_width.restype = ctypes.c_int
    lib.mds_oracle_max_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def sum_of_increments(delta, x, precision, truncated, fp32=False) -> float:
    lib = load()
    delta, pd = _c(delta)
    x, px = _c(x)
    n = x.shape[0]
    p = x.size // n
    fn = lib.mds_oracle_sum_of_increments_f32 if fp32 else lib.mds_oracle_sum_of_increments_f64
    return fn(pd, px, n, p, float(precision), int(bool(truncated)))


def gradient(delta, x, precision, truncated, fp32=False) -> np.ndarray:
    lib = load()
    delta, pd = _c(delta)
    x, px = _c(x)
    n = x.shape[0]
    p = x.size // n
    out = np.empty(n * p)
    fn = lib.mds_oracle_gradient_f32 if fp32 else lib.mds_oracle_gradient_f64
    fn(pd, px, n, p, float(precision), int(bool(truncated)), out.ctypes.data_as(_dp))
    return out.reshape(n, p)


def rows_sum_of_increments(delta, x, precision, truncated, rows) -> float:
    """Serial FP64 sum over the first `rows` rows (timing samples): with
    rows < N the raw row-sum total (unhalved, unscaled), as
    simd_sum_of_increments returns for a sample."""
    lib = load()
    delta, pd = _c(delta)
    x, px = _c(x)
    n = x.shape[0]
    return lib.mds_oracle_rows_sum_of_increments_f64(pd, px, n, x.size // n, float(precision),
                                                     int(bool(truncated)), int(rows))


def rows_gradient(delta, x, precision, truncated, rows) -> np.ndarray:
    """Serial FP64 gradient rows [0, rows) (timing samples; other rows 0)."""
    lib = load()
    delta, pd = _c(delta)
    x, px = _c(x)
    n = x.shape[0]
    p = x.size // n
    out = np.empty(n * p)
    lib.mds_oracle_rows_gradient_f64(pd, px, n, p, float(precision), int(bool(truncated)), int(rows),
                                     out.ctypes.data_as(_dp))
    return out.reshape(n, p)


def log_likelihood(sum_of_increments_value, n, precision) -> float:
    """R/mds.R:24-28."""
    return load().mds_oracle_log_likelihood(float(sum_of_increments_value), int(n), float(precision))


def simd_sum_of_increments(delta, x, precision, truncated, threads=0, rows=0) -> float:
    lib = load()
    delta, pd = _c(delta)
    x, px = _c(x)
    n = x.shape[0]
    p = x.size // n
    return lib.mds_oracle_simd_sum_of_increments(pd, px, n, p, float(precision), int(bool(truncated)),
                                                 int(threads), int(rows))


def simd_gradient(delta, x, precision, truncated, threads=0, rows=0) -> np.ndarray:
    lib = load()
    delta, pd = _c(delta)
    x, px = _c(x)
    n = x.shape[0]
    p = x.size // n
    out = np.zeros(n * p)
    lib.mds_oracle_simd_gradient(pd, px, n, p, float(precision), int(bool(truncated)), int(threads),
                                 int(rows), out.ctypes.data_as(_dp))
    return out.reshape(n, p)


def simd_width() -> int:
    return load().mds_oracle_simd_width()


def max_threads() -> int:
    return load().mds_oracle_max_threads()


# --------------------------------------------------------------------------
# High-precision closed form (mpmath), independent of the engine arithmetic.
# --------------------------------------------------------------------------

def closed_form_loglik_mp(delta, x, precision, truncated, dps=50, row_only=None):
    """log-likelihood as R's computeLoglikelihood (R/Untitled.R:16-52):

        sum_{i<j} [ log dnorm(delta_ij; d_ij, 1/sqrt(tau)) - T log pnorm(d_ij sqrt(tau)) ]

    over pairs whose observation is not NaN (the R code has no missing-data
    rule; the engine skips NaN pairs).  ``x`` may hold mpmath numbers.  With
    ``row_only = i`` the sum runs over the ordered pairs (i, j) of row i only
    (no 1/2), the quantity whose x_i-gradient the reference's row-wise
    gradient loop computes for asymmetric data.
    """
    import mpmath as mp

    with mp.workdps(dps):
        n = len(delta)
        tau = mp.mpf(precision)
        sd = 1 / mp.sqrt(tau)
        log2pi = mp.log(2 * mp.pi)
        total = mp.mpf(0)

        def term(o, i, j):
            d = mp.sqrt(sum((mp.mpf(x[i][k]) - mp.mpf(x[j][k])) ** 2 for k in range(len(x[i]))))
            o = mp.mpf(o)
            v = -mp.mpf(0.5) * log2pi - mp.log(sd) - (o - d) ** 2 / (2 * sd * sd)
            if truncated:
                v -= mp.log(mp.ncdf(d / sd))
            return v

        if row_only is not None:
            i = row_only
            for j in range(n):
                o = float(delta[i][j])
                if j != i and not np.isnan(o):
                    total += term(o, i, j)
            return total
        for i in range(n):
            for j in range(i + 1, n):
                o = float(delta[i][j])
                if not np.isnan(o):
                    total += term(o, i, j)
        return total


def closed_form_gradient_fd_mp(delta, x, precision, truncated, dps=50, h="1e-15", row_only=False):
    """Central finite differences of ``closed_form_loglik_mp`` at ``dps``
    digits (truncation error O(h^2) ~ 1e-30): an independent check of the
    analytic gradient.  With ``row_only`` the x_i derivative is taken of the
    row-i sum, which is what the reference gradient computes row by row."""
    import mpmath as mp

    with mp.workdps(dps):
        n, p = np.asarray(x).shape
        hh = mp.mpf(h)
        g = np.zeros((n, p))
        for i in range(n):
            for k in range(p):
                xp = [[mp.mpf(float(v)) for v in row] for row in x]
                xm = [[mp.mpf(float(v)) for v in row] for row in x]
                xp[i][k] += hh
                xm[i][k] -= hh
                r = i if row_only else None
                fp = closed_form_loglik_mp(delta, xp, precision, truncated, dps, row_only=r)
                fm = closed_form_loglik_mp(delta, xm, precision, truncated, dps, row_only=r)
                g[i, k] = float((fp - fm) / (2 * hh))
        return g
