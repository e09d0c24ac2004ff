"""ctypes binding of libmds_hip.so, shaped like the reference engine interface.

``MultiDimensionalScaling`` mirrors ``mds::AbstractMultiDimensionalScaling``
(src/AbstractMultiDimensionalScaling.hpp:38-77) method for method, and
``factory`` mirrors ``mds::factory`` (src/factory.cpp:32-118) for the GPU
route.  Every call goes through the C ABI declared in include/mds_hip.h; a
non-zero status raises ``MdsError`` with the engine's message.  There is no
CPU fallback: without the built library or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libmds_hip.so"


class Flags:
    """Bit values of ``mds::Flags`` (src/flags.h:6-14) plus the HIP alias."""

    FLOAT = 1 << 2
    TBB = 1 << 3
    OPENCL = 1 << 4
    LEFT_TRUNCATION = 1 << 5
    SSE = 1 << 7
    AVX = 1 << 8
    AVX512 = 1 << 9
    HIP = 1 << 10


class MdsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


_lib = None

_c_double_p = ctypes.POINTER(ctypes.c_double)


def library_path() -> Path:
    return Path(os.environ.get("MDS_HIP_LIBRARY", str(_LIB_PATH)))


def load_library() -> ctypes.CDLL:
    """Loads libmds_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with `python -m massivemds_amd.build`")
    lib = ctypes.CDLL(str(path))
    H = ctypes.c_void_p
    sig = {
        "mds_hip_max_dimension": ([], ctypes.c_int),
        "mds_hip_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "mds_hip_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_int,
                            ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "mds_hip_destroy": ([H], ctypes.c_int),
        "mds_hip_last_error": ([H], ctypes.c_char_p),
        "mds_hip_update_locations": ([H, ctypes.c_int, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_get_sum_of_increments": ([H, _c_double_p], ctypes.c_int),
        "mds_hip_get_log_likelihood_gradient": ([H, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_store_state": ([H], ctypes.c_int),
        "mds_hip_restore_state": ([H], ctypes.c_int),
        "mds_hip_accept_state": ([H], ctypes.c_int),
        "mds_hip_set_pairwise_data": ([H, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_set_parameters": ([H, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_make_dirty": ([H], ctypes.c_int),
        "mds_hip_get_internal_dimension": ([H, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "mds_hip_evaluate_async": ([H, ctypes.c_int], ctypes.c_int),
        "mds_hip_synchronize": ([H], ctypes.c_int),
        "mds_hip_evaluate_device_noalloc": ([H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "mds_hip_set_timing": ([H, ctypes.c_int], ctypes.c_int),
        "mds_hip_get_timing": ([H, _c_double_p, _c_double_p, ctypes.POINTER(ctypes.c_longlong)],
                               ctypes.c_int),
        "mds_hip_get_layout": ([H, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong),
                                ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "mds_hip_get_counters": ([H, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
                                 ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def device_count() -> int:
    lib = load_library()
    c = ctypes.c_int(0)
    lib.mds_hip_device_count(ctypes.byref(c))
    return c.value


def _as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_c_double_p)


class MultiDimensionalScaling:
    """GPU engine instance (src/AbstractMultiDimensionalScaling.hpp:38-73)."""

    def __init__(self, embeddingDimension: int, locationCount: int, flags: int, device: int = 0):
        self._lib = load_library()
        h = ctypes.c_void_p()
        st = self._lib.mds_hip_create(int(embeddingDimension), int(locationCount), int(flags),
                                      int(device), ctypes.byref(h))
        if st != 0:
            raise MdsError(st, self._lib.mds_hip_last_error(None).decode())
        self._h = h
        self.embeddingDimension = int(embeddingDimension)
        self.locationCount = int(locationCount)
        self.flags = int(flags)

    # -- helpers -------------------------------------------------------------
    def _check(self, st: int) -> None:
        if st != 0:
            raise MdsError(st, self._lib.mds_hip_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.mds_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference interface ----------------------------------------------------
    def updateLocations(self, locationIndex: int, location) -> None:
        a = _as_f64(location)
        self._check(self._lib.mds_hip_update_locations(self._h, int(locationIndex), _ptr(a), a.size))

    def getSumOfIncrements(self) -> float:
        out = ctypes.c_double()
        self._check(self._lib.mds_hip_get_sum_of_increments(self._h, ctypes.byref(out)))
        return out.value

    def getLogLikelihoodGradient(self, length: int | None = None) -> np.ndarray:
        n = self.locationCount * self.embeddingDimension if length is None else int(length)
        out = np.empty(n)
        self._check(self._lib.mds_hip_get_log_likelihood_gradient(self._h, _ptr(out), n))
        return out

    def storeState(self) -> None:
        self._check(self._lib.mds_hip_store_state(self._h))

    def restoreState(self) -> None:
        self._check(self._lib.mds_hip_restore_state(self._h))

    def acceptState(self) -> None:
        self._check(self._lib.mds_hip_accept_state(self._h))

    def setPairwiseData(self, data) -> None:
        a = _as_f64(data)
        self._check(self._lib.mds_hip_set_pairwise_data(self._h, _ptr(a), a.size))

    def setParameters(self, parameters) -> None:
        a = _as_f64(parameters)
        self._check(self._lib.mds_hip_set_parameters(self._h, _ptr(a), a.size))

    def makeDirty(self) -> None:
        self._check(self._lib.mds_hip_make_dirty(self._h))

    def getInternalDimension(self) -> int:
        out = ctypes.c_int()
        self._check(self._lib.mds_hip_get_internal_dimension(self._h, ctypes.byref(out)))
        return out.value

    # -- device-resident extensions ---------------------------------------------
    def evaluate_async(self, gradient: bool = True) -> None:
        self._check(self._lib.mds_hip_evaluate_async(self._h, 1 if gradient else 0))

    def synchronize(self) -> None:
        self._check(self._lib.mds_hip_synchronize(self._h))

    def evaluate_device_noalloc(self, locations_ptr: int | None, sum_ptr: int, gradient_ptr: int | None) -> None:
        """mds_hip_evaluate_device_noalloc: device addresses of the N x P FP64
        locations (None: the current ones), of one FP64 for the sum and of
        the N x P FP64 gradient (None: likelihood only)."""
        self._check(self._lib.mds_hip_evaluate_device_noalloc(self._h, locations_ptr, sum_ptr, gradient_ptr))

    def set_timing(self, enable: bool) -> None:
        self._check(self._lib.mds_hip_set_timing(self._h, 1 if enable else 0))

    def get_timing(self):
        """(sum of tile-kernel ms, sum of whole-evaluation ms, count)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        self._check(self._lib.mds_hip_get_timing(self._h, ctypes.byref(a), ctypes.byref(b),
                                                 ctypes.byref(c)))
        return a.value, b.value, c.value

    def counters(self):
        """(full O(N^2) evaluations launched, sums from O(N) single-row updates)."""
        f, r = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self._lib.mds_hip_get_counters(self._h, ctypes.byref(f), ctypes.byref(r)))
        return f.value, r.value

    def layout(self):
        """(symmetric, packed observation bytes, tile count)."""
        s, b, t = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self._lib.mds_hip_get_layout(self._h, ctypes.byref(s), ctypes.byref(b),
                                                 ctypes.byref(t)))
        return bool(s.value), b.value, t.value


def factory(dim1: int, dim2: int, flags: int, device: int = 0, threads: int = 0):
    """``mds::factory(P, N, flags, device, threads)`` for the GPU route.

    The OPENCL bit (src/flags.h:9) or the HIP alias selects this engine, with
    FLOAT choosing the compensated FP32 kernels (src/factory.cpp:49-80).  The
    reference's CPU engines are not part of this library, so CPU flag sets
    raise ``ValueError``.
    """
    if not flags & (Flags.OPENCL | Flags.HIP):
        raise ValueError("CPU engines are not provided by libmds_hip; set Flags.OPENCL or Flags.HIP")
    return MultiDimensionalScaling(dim1, dim2, flags, device)
