"""Sizes past the reference's 32-bit indexing: N = 46,400 > 46,341, where
N (N - 1) and i * N + j overflow an int (src/AbstractMultiDimensionalScaling.hpp:42,
src/NewMultiDimensionalScaling.hpp:516).  The oracle cannot finish at this size
in a test, so the checks are size-independent properties:

* a closed form: constant observations delta = 1.5 with one missing row and
  column, every location at the origin, no truncation: every pair term is
  r^2 = 2.25, a multiple of 1/4, so every partial sum is exact and the sum of
  increments must equal 0.5 * tau * 0.5 * (N - 1) (N - 2) * 2.25 bit for bit;
* on random locations: the gradient rows sum to zero (translation invariance),
  the likelihood is unchanged by a translation of every location, and the FP64
  and FP32 engines agree.

Host memory: the N x N input (17 GB) and the packing buffer (17 GB); device:
8.6 GB of packed observations."""
import numpy as np
import pytest

from massivemds_amd import datagen
from massivemds_amd.engine import Flags, MultiDimensionalScaling

pytestmark = pytest.mark.gpu

N = 46400


@pytest.fixture(scope="module")
def delta():
    print(f"max-size test: N = {N}, building the {N} x {N} input", flush=True)
    d = np.full((N, N), 1.5)
    np.fill_diagonal(d, 0.0)
    d[5, :] = np.nan
    d[:, 5] = np.nan
    return d


def test_closed_form_sum_past_int_range(delta):
    tau = 2.0
    e = MultiDimensionalScaling(2, N, Flags.OPENCL, 0)
    try:
        e.setPairwiseData(delta)
        e.updateLocations(-1, np.zeros((N, 2)))
        e.setParameters([tau])
        got = e.getSumOfIncrements()
    finally:
        e.close()
    expected = 0.5 * tau * 0.5 * float((N - 1) * (N - 2)) * 2.25
    assert got == expected, (got, expected)


def test_invariants_past_int_range(delta):
    tau = 1.3
    x = datagen.locations(N, 2)
    shift = np.array([0.75, -1.25])
    out = {}
    for name, flags in (("f64", Flags.OPENCL | Flags.LEFT_TRUNCATION),
                        ("f32", Flags.OPENCL | Flags.LEFT_TRUNCATION | Flags.FLOAT)):
        e = MultiDimensionalScaling(2, N, flags, 0)
        try:
            e.setPairwiseData(delta)
            e.setParameters([tau])
            e.updateLocations(-1, x)
            s = e.getSumOfIncrements()
            g = e.getLogLikelihoodGradient().reshape(N, 2)
            e.updateLocations(-1, x + shift)
            s_shift = e.getSumOfIncrements()
        finally:
            e.close()
        out[name] = (s, g, s_shift)
    s, g, s_shift = out["f64"]
    assert np.isfinite(s) and np.all(np.isfinite(g))
    scale = np.max(np.abs(g))
    assert np.all(np.abs(g.sum(axis=0)) <= 1e-9 * N * scale)
    assert abs(s_shift - s) <= 1e-12 * abs(s)
    assert np.all(g[5] == 0.0)  # row 5 has no observed pair
    s32, g32, _ = out["f32"]
    assert abs(s32 - s) <= 1e-6 * abs(s)
    assert np.max(np.abs(g32 - g)) <= 1e-5 * scale
