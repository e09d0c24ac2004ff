"""Accuracy of the kernels' fused truncation arithmetic, checked on the CPU.

libmds_mathcheck.so compiles mds_trunc_math.hpp for the host (the hardware
reciprocal-square-root seed replaced by one of the same accuracy).  Its d,
1/d, log Phi(z) and phi(z) / (Phi(z) sqrt tau) are compared with mpmath at
40 digits over the whole range of z that matters (and beyond, where the
terms vanish).

Budget: the FP64 engine must match the reference to 1e-12 relative on sums
of up to ~1e8 terms.  d and 1/d are checked relative; the two table terms in
absolute error, as they enter O(1) sums (log Phi) and the difference
r - phi / (Phi sqrt tau).
"""
import ctypes

import mpmath as mp
import numpy as np
import pytest

from massivemds_amd import build as hipbuild

_dp = ctypes.POINTER(ctypes.c_double)


def _lib():
    if not hipbuild.MATHCHECK_LIB.exists():
        hipbuild.build()
    lib = ctypes.CDLL(str(hipbuild.MATHCHECK_LIB))
    for name in ("mds_mathcheck_f64", "mds_mathcheck_f32"):
        fn = getattr(lib, name)
        fn.argtypes = [_dp, ctypes.c_longlong, ctypes.c_double, _dp, _dp, _dp, _dp]
        fn.restype = None
    return lib


def _eval(fp32, d2, tau):
    lib = _lib()
    d2 = np.ascontiguousarray(d2, dtype=np.float64)
    outs = [np.empty_like(d2) for _ in range(4)]
    fn = lib.mds_mathcheck_f32 if fp32 else lib.mds_mathcheck_f64
    fn(d2.ctypes.data_as(_dp), d2.size, tau, *[o.ctypes.data_as(_dp) for o in outs])
    return outs


def _exact(d2, tau):
    mp.mp.dps = 40
    d = mp.sqrt(mp.mpf(d2))
    z = d * mp.sqrt(tau)
    phi_cdf = mp.ncdf(z)
    log_phi = mp.log(phi_cdf)
    grad = mp.npdf(z) / (phi_cdf * mp.sqrt(tau))
    return float(d), float(log_phi), float(grad)


def _grid(tau):
    z = np.concatenate([np.linspace(0.0, 1.0, 101), np.linspace(1.0, 9.0, 401), np.linspace(9.0, 40.0, 63),
                        np.array([1e-8, 1e-4, 0.3, 2.5, 12.0, 11.999, 12.001, 37.0, 100.0])])
    return (z / np.sqrt(tau)) ** 2


@pytest.mark.parametrize("tau", [0.5, 1.0, 2.0, 3.7])
def test_fp64_terms(tau):
    """One Newton step for d and 1/d (relative error ~1.5 e0^2 from the
    ~2^-26 seed), log Phi and phi / Phi from the interval table."""
    d2 = _grid(tau)
    d, invd, log_phi, grad = _eval(False, d2, tau)
    for k, v in enumerate(d2):
        ed, elog, egrad = _exact(v, tau)
        assert abs(d[k] - ed) <= 8e-16 * ed, (v, d[k], ed)
        if v > 0:
            assert abs(invd[k] * ed - 1) <= 8e-16
        assert abs(log_phi[k] - elog) <= 4e-16, (v, log_phi[k], elog)
        assert abs(grad[k] - egrad) <= 2e-14 / np.sqrt(tau), (v, grad[k], egrad)


def test_fp64_table_edges():
    """Interval edges (t = 0 of every row), the clamp to the zero row, and
    distances that are not finite."""
    tau = 2.0
    z = np.concatenate([np.arange(0, 641) / 64.0, [10.0 - 1e-12, 10.0, 10.0 + 1e-12, 1e3]])
    d2 = (z / np.sqrt(tau)) ** 2
    d, invd, log_phi, grad = _eval(False, d2, tau)
    for k, v in enumerate(d2):
        _, elog, egrad = _exact(v, tau)
        assert abs(log_phi[k] - elog) <= 4e-16 and abs(grad[k] - egrad) <= 2e-14, (z[k], log_phi[k], elog)
    d, invd, log_phi, grad = _eval(False, np.array([np.inf, np.nan]), tau)
    assert np.all(log_phi == 0.0) and np.all(grad == 0.0)
    assert np.isnan(d[1]) and np.isnan(invd[1])  # NaN propagates through d (and so the residual)
    assert d[0] == np.inf  # an infinite distance stays infinite (the residual is -inf)
    # the guard on 1/d leaves tiny but representable distances exact
    d, invd, _, _ = _eval(False, np.array([1e-300, 2.0 ** -998]), tau)
    assert abs(d[0] / 1e-150 - 1) <= 8e-16 and abs(d[1] / 2.0 ** -499 - 1) <= 8e-16


def test_fp64_zero_distance():
    for fp32 in (False, True):
        d, invd, log_phi, grad = _eval(fp32, np.array([0.0]), 2.0)
        # 1/d is not finite at d = 0 (NaN after the FP64 Newton step, inf
        # from the FP32 seed): the gradient term (r / d) (x_i - x_j) is NaN
        # there, as in the reference (inf * 0)
        assert d[0] == 0.0 and not np.isfinite(invd[0])
        assert abs(log_phi[0] - np.log(0.5)) <= (1e-7 if fp32 else 5e-16)
        assert abs(grad[0] - 2 * 0.3989422804014327 / np.sqrt(2.0)) <= (2e-7 if fp32 else 2e-14)


@pytest.mark.parametrize("tau", [0.5, 2.0])
def test_fp32_terms(tau):
    d2 = _grid(tau)
    d, invd, log_phi, grad = _eval(True, d2, tau)
    for k, v in enumerate(d2):
        ed, elog, egrad = _exact(np.float64(np.float32(v)), tau)
        assert abs(d[k] - ed) <= 3e-7 * ed
        assert abs(log_phi[k] - elog) <= 2e-7, (v, log_phi[k], elog)
        # table error ~1e-7 plus the rounding of z in float (~1e-7 z relative)
        z = np.sqrt(v * tau)
        assert abs(grad[k] - egrad) <= (2e-7 + 1e-7 * z) / np.sqrt(tau), (v, grad[k], egrad)
