"""Batched engine: many independent datasets of their own N and P evaluated
in one persistent launch (C ABI: mds_hip_batch_* in include/mds_hip.h).

Each dataset keeps the single-instance semantics: its own observations,
locations and precision; `sums()` returns every dataset's
getSumOfIncrements() and `gradient(k)` the N_k x P_k log-likelihood gradient.
Entries can share another entry's observations (`sharePairwiseData`), so
that several chains of one dataset run side by side, and `run_hmc` runs the
device HMC sampler on every entry at once, one launch per leapfrog step for
all chains.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .engine import MdsError, _as_f64, _c_double_p, _ptr, load_library

_bound = False


def _bind(lib):
    global _bound
    if _bound:
        return lib
    H = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "mds_hip_batch_create": ([ctypes.c_int, ip, ip, ctypes.c_long, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "mds_hip_batch_destroy": ([H], ctypes.c_int),
        "mds_hip_batch_last_error": ([H], ctypes.c_char_p),
        "mds_hip_batch_set_pairwise_data": ([H, ctypes.c_int, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_batch_update_locations": ([H, ctypes.c_int, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_batch_set_parameters": ([H, ctypes.c_int, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_batch_evaluate": ([H], ctypes.c_int),
        "mds_hip_batch_get_sums_of_increments": ([H, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_batch_get_gradient": ([H, ctypes.c_int, _c_double_p, ctypes.c_size_t], ctypes.c_int),
        "mds_hip_batch_autotune": ([H, ctypes.c_char_p, ctypes.c_int, _c_double_p], ctypes.c_int),
        "mds_hip_batch_set_config": ([H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "mds_hip_batch_get_config": ([H, ip, ip, ip, ip], ctypes.c_int),
        "mds_hip_batch_set_timing": ([H, ctypes.c_int], ctypes.c_int),
        "mds_hip_batch_get_timing": ([H, _c_double_p, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "mds_hip_batch_share_pairwise_data": ([H, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "mds_hip_batch_hmc_run": ([H, ctypes.c_void_p, _c_double_p, ctypes.c_size_t, _c_double_p, ctypes.c_size_t,
                                   _c_double_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _bound = True
    return lib


class MultiDimensionalScalingBatch:
    def __init__(self, embeddingDimensions, locationCounts, flags: int, device: int = 0):
        self._lib = _bind(load_library())
        dims = np.ascontiguousarray(embeddingDimensions, dtype=np.int32)
        ns = np.ascontiguousarray(locationCounts, dtype=np.int32)
        if dims.size != ns.size:
            raise ValueError("one embedding dimension per dataset")
        h = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int)
        st = self._lib.mds_hip_batch_create(int(dims.size), dims.ctypes.data_as(ip), ns.ctypes.data_as(ip),
                                            int(flags), int(device), ctypes.byref(h))
        if st != 0:
            raise MdsError(st, self._lib.mds_hip_batch_last_error(None).decode())
        self._h = h
        self.dims = [int(v) for v in dims]
        self.ns = [int(v) for v in ns]

    def _check(self, st):
        if st != 0:
            raise MdsError(st, self._lib.mds_hip_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mds_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setPairwiseData(self, index: int, data) -> None:
        a = _as_f64(data)
        self._check(self._lib.mds_hip_batch_set_pairwise_data(self._h, int(index), _ptr(a), a.size))

    def updateLocations(self, index: int, locations) -> None:
        a = _as_f64(locations)
        self._check(self._lib.mds_hip_batch_update_locations(self._h, int(index), _ptr(a), a.size))

    def setParameters(self, index: int, parameters) -> None:
        a = _as_f64(parameters)
        self._check(self._lib.mds_hip_batch_set_parameters(self._h, int(index), _ptr(a), a.size))

    def evaluate(self) -> None:
        self._check(self._lib.mds_hip_batch_evaluate(self._h))

    def sums(self) -> np.ndarray:
        out = np.empty(len(self.ns))
        self._check(self._lib.mds_hip_batch_get_sums_of_increments(self._h, _ptr(out), out.size))
        return out

    def gradient(self, index: int) -> np.ndarray:
        n, p = self.ns[index], self.dims[index]
        out = np.empty(n * p)
        self._check(self._lib.mds_hip_batch_get_gradient(self._h, int(index), _ptr(out), out.size))
        return out.reshape(n, p)

    def autotune(self, cache: str | None = None, force: bool = False) -> float:
        ms = ctypes.c_double()
        path = cache.encode() if cache else None
        self._check(self._lib.mds_hip_batch_autotune(self._h, path, int(force), ctypes.byref(ms)))
        return ms.value

    def set_config(self, wavesPerBlock: int, blocksPerCU: int, tilesPerItem: int, strategy: int) -> None:
        self._check(self._lib.mds_hip_batch_set_config(self._h, wavesPerBlock, blocksPerCU, tilesPerItem, strategy))

    def get_config(self):
        vals = [ctypes.c_int() for _ in range(4)]
        self._check(self._lib.mds_hip_batch_get_config(self._h, *[ctypes.byref(v) for v in vals]))
        return tuple(v.value for v in vals)

    def set_timing(self, enable: bool) -> None:
        self._check(self._lib.mds_hip_batch_set_timing(self._h, int(bool(enable))))

    def get_timing(self):
        ms, c = ctypes.c_double(), ctypes.c_longlong()
        self._check(self._lib.mds_hip_batch_get_timing(self._h, ctypes.byref(ms), ctypes.byref(c)))
        return ms.value, c.value

    def sharePairwiseData(self, index: int, source: int) -> None:
        """Entry `index` reads entry `source`'s observations (same N)."""
        self._check(self._lib.mds_hip_batch_share_pairwise_data(self._h, int(index), int(source)))

    def run_hmc(self, initials, options):
        """The device HMC sampler on every entry at once (one chain per
        entry; chain k uses seed options.seed + k).  `initials`: one N_k x P_k
        array per entry.  Returns (traces, final locations, stats), one per
        entry, with the single-chain sampler's trace columns."""
        from .hmc import TRACE_COLUMNS, HmcStats

        if len(initials) != len(self.ns):
            raise ValueError("one initial location array per entry")
        x0 = np.concatenate([np.ascontiguousarray(a, dtype=np.float64).ravel() for a in initials])
        kept = options.iterations - options.burnIn
        count = len(self.ns)
        traces = np.empty(count * kept * len(TRACE_COLUMNS))
        finals = np.empty(x0.size)
        stats = (HmcStats * count)()
        self._check(self._lib.mds_hip_batch_hmc_run(self._h, ctypes.addressof(options), _ptr(x0), x0.size,
                                                    _ptr(traces), traces.size, _ptr(finals), finals.size,
                                                    ctypes.addressof(stats), count))
        traces = traces.reshape(count, kept, len(TRACE_COLUMNS))
        out, k = [], 0
        for n, p in zip(self.ns, self.dims):
            out.append(finals[k:k + n * p].reshape(n, p))
            k += n * p
        return list(traces), out, list(stats)
