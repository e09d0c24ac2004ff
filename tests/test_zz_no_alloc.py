"""GPU test of the allocation-free evaluation entry point
(mds_hip_evaluate_device_noalloc, through the C ABI): device-resident
locations in, the sum and gradient out into caller device buffers, the
partials in the library's static scratch array.  Each call is compared with
mds_hip_evaluate_async at the same locations (the same kernels in the same
order: bitwise equal) and with the serial FP64 oracle (FP64 engine 1e-12;
FP32 engine 2e-6 on the sum, 2e-5 of max|g| on the gradient); two calls in a
row share the array, the second on a smaller instance, then a third on the
first instance again."""
import ctypes

import numpy as np
import pytest

from massivemds_amd import datagen
from massivemds_amd.engine import Flags, MdsError, MultiDimensionalScaling, load_library

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Device:
    """Device buffers through the HIP runtime the library itself links."""

    def __init__(self):
        load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.held = []

    def alloc(self, count):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), 8 * count) == 0
        self.held.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        p = self.alloc(a.size)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, _H2D) == 0
        return p

    def download(self, p, count):
        out = np.empty(count)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.hip.hipFree(p)
        self.held = []


def _instance(n, p, truncated, fp32, tau, seed):
    flags = Flags.OPENCL | (Flags.LEFT_TRUNCATION if truncated else 0) | (Flags.FLOAT if fp32 else 0)
    delta = datagen.euclid10(n, seed, 0.1, 0.05)
    e = MultiDimensionalScaling(p, n, flags, 0)
    e.setPairwiseData(delta)
    e.setParameters([tau])
    return e, delta


def _existing(e, x):
    """The existing entry point: host locations, mds_hip_evaluate_async,
    the sum and gradient read back."""
    e.updateLocations(-1, x)
    e.evaluate_async(True)
    return e.getSumOfIncrements(), e.getLogLikelihoodGradient().reshape(x.shape)


def _noalloc(dev, e, x):
    n, p = x.shape
    dx = dev.upload(x)
    ds = dev.alloc(1)
    dg = dev.alloc(n * p)
    e.evaluate_device_noalloc(dx, ds, dg)
    return dev.download(ds, 1)[0], dev.download(dg, n * p).reshape(n, p)


def _check_case(dev, oracle, e, delta, x, tau, truncated, fp32):
    print(f"noalloc n={x.shape[0]} p={x.shape[1]} T={truncated} fp32={fp32}", flush=True)
    s_new, g_new = _noalloc(dev, e, x)
    s_old, g_old = _existing(e, x)
    assert s_new == s_old
    assert np.array_equal(g_new, g_old)
    s_ref = oracle.sum_of_increments(delta, x, tau, truncated)
    g_ref = oracle.gradient(delta, x, tau, truncated)
    tol_s, tol_g = (2e-6, 2e-5) if fp32 else (1e-12, 1e-12)
    assert abs(s_new - s_ref) <= tol_s * abs(s_ref), (s_new, s_ref)
    assert np.max(np.abs(g_new - g_ref)) <= tol_g * np.max(np.abs(g_ref))


def test_noalloc_matches_existing_entry_point(oracle):
    dev = _Device()
    try:
        tau_a, tau_b = 2.0, 0.7
        ea, da = _instance(3000, 2, True, False, tau_a, 71)
        eb, db = _instance(700, 3, False, True, tau_b, 72)
        xa = datagen.locations(3000, 2, 73)
        xb = datagen.locations(700, 3, 74)
        _check_case(dev, oracle, ea, da, xa, tau_a, True, False)
        _check_case(dev, oracle, eb, db, xb, tau_b, False, True)   # smaller, after the larger one
        _check_case(dev, oracle, ea, da, xa + 0.05, tau_a, True, False)
        # current locations, likelihood alone
        ds = dev.alloc(1)
        eb.evaluate_device_noalloc(None, ds, None)
        s_cur = dev.download(ds, 1)[0]
        assert s_cur == eb.getSumOfIncrements()
        ea.close()
        eb.close()
    finally:
        dev.free()


def test_noalloc_locations_leave_a_stored_state_in_place(oracle):
    """storeState, then new locations through the allocation-free entry, then
    restoreState: the engine holds the stored locations again (evaluated
    afresh: bitwise the existing entry point at those locations)."""
    dev = _Device()
    try:
        tau = 1.1
        e, delta = _instance(500, 2, True, False, tau, 77)
        x1 = datagen.locations(500, 2, 78)
        x2 = datagen.locations(500, 2, 79)
        e.updateLocations(-1, x1)
        e.getSumOfIncrements()
        e.storeState()
        s2, _ = _noalloc(dev, e, x2)
        e.restoreState()
        e.makeDirty()
        s1 = e.getSumOfIncrements()
        g1 = e.getLogLikelihoodGradient()
        e.makeDirty()
        s1_again, g1_again = _existing(e, x1)
        assert s1 == s1_again and np.array_equal(g1, g1_again.ravel())
        ref1 = oracle.sum_of_increments(delta, x1, tau, True)
        ref2 = oracle.sum_of_increments(delta, x2, tau, True)
        assert abs(s1 - ref1) <= 1e-12 * abs(ref1)
        assert abs(s2 - ref2) <= 1e-12 * abs(ref2)
        e.close()
    finally:
        dev.free()


def test_noalloc_refuses_bad_buffers_and_sizes():
    dev = _Device()
    try:
        e, _ = _instance(200, 2, True, False, 1.0, 75)
        x = datagen.locations(200, 2, 76)
        dx = dev.upload(x)
        ds = dev.alloc(1)
        host = np.zeros(400)
        with pytest.raises(MdsError) as err:  # a host buffer never reaches a kernel
            e.evaluate_device_noalloc(dx, host.ctypes.data, None)
        assert err.value.status == 1
        block = dev.alloc(1 << 19)  # 4 MiB
        with pytest.raises(MdsError) as err:  # N x P doubles from 8 bytes before the allocation's end
            e.evaluate_device_noalloc(dx, ds, block + (1 << 22) - 8)
        assert err.value.status == 1
        with pytest.raises(MdsError) as err:
            e.evaluate_device_noalloc(dx, None, None)
        assert err.value.status == 1
        # a valid call still works after the refusals
        e.evaluate_device_noalloc(dx, ds, None)
        e.close()
        # partials beyond the scratch array: refused before anything runs
        big = MultiDimensionalScaling(2, 20000, Flags.OPENCL | Flags.LEFT_TRUNCATION, 0)
        g = dev.alloc(40000)
        with pytest.raises(MdsError) as err:
            big.evaluate_device_noalloc(None, ds, g)
        assert err.value.status == 1 and "scratch" in str(err.value)
        big.close()
    finally:
        dev.free()
