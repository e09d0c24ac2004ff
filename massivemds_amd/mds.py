"""Python counterpart of the reference's R engine API (R/mds.R, src/MassiveMDS.cpp).

The R package wraps an engine in a list and exposes ``createEngine``,
``setPairwiseData``, ``updateLocations``, ``setPrecision``,
``getLogLikelihood`` and ``getGradient``.  The same functions here drive the
HIP engine through ``massivemds_amd.engine`` with the same argument meaning,
checks and results.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .engine import Flags, MultiDimensionalScaling, factory


@dataclass
class Engine:
    """The R engine list (src/MassiveMDS.cpp:101-111)."""

    engine: MultiDimensionalScaling
    embeddingDimension: int
    locationCount: int
    truncation: bool
    threads: int
    deviceNumber: int
    flags: int
    dataInitialized: bool = False
    locationsInitialized: bool = False
    precision: float | None = None
    extra: dict = field(default_factory=dict)


def createEngine(embeddingDimension: int, locationCount: int, truncation: bool, tbb: int = 0,
                 simd: int = 0, gpu: int = 1, single: bool = False) -> Engine:
    """src/MassiveMDS.cpp:60-114.  ``gpu > 0`` sets the GPU flag and selects
    HIP device ``gpu - 1`` (counting GPUs from 1, as the reference's CLI does
    with ``--gpu k``, src/benchmark.cpp:89-93; the R wrapper passes ``gpu``
    itself, which on a one-GPU machine only reached device 0 through the
    engine's out-of-range fallback); ``single`` selects the FP32 engine.  CPU
    engines (gpu = 0) belong to the reference library and are not provided
    here."""
    flags = 0
    if truncation:
        flags |= Flags.LEFT_TRUNCATION
    if gpu <= 0:
        raise ValueError("this library provides the GPU engine only: call createEngine(..., gpu=1)")
    flags |= Flags.OPENCL
    device = int(gpu) - 1
    if single:
        flags |= Flags.FLOAT
    eng = factory(embeddingDimension, locationCount, flags, device, 0)
    return Engine(eng, int(embeddingDimension), int(locationCount), bool(truncation), 0, device, flags)


def setPairwiseData(engine: Engine, data) -> Engine:
    """R/mds.R:84-92.  As R's ``as.vector`` on a matrix, the matrix is passed
    column-major (identical to row-major for symmetric data)."""
    data = np.asarray(data, dtype=np.float64)
    flat = data.ravel(order="F") if data.ndim == 2 else data.ravel()
    if flat.size != engine.locationCount * engine.locationCount:
        raise ValueError("Invalid data size")
    engine.engine.setPairwiseData(flat)
    engine.dataInitialized = True
    return engine


def updateLocations(engine: Engine, locations) -> Engine:
    """R/mds.R:103-111: N x P locations sent row-major."""
    loc = np.asarray(locations, dtype=np.float64)
    flat = loc.reshape(engine.locationCount, -1).ravel() if loc.ndim == 2 else loc.ravel()
    if flat.size != engine.locationCount * engine.embeddingDimension:
        raise ValueError("Invalid data size")
    engine.engine.updateLocations(-1, flat)
    engine.locationsInitialized = True
    return engine


def setPrecision(engine: Engine, precision: float) -> Engine:
    """R/mds.R:68-72 (setParameters with the precision)."""
    engine.engine.setParameters([float(precision)])
    engine.precision = float(precision)
    return engine


def _check_ready(engine: Engine) -> None:
    if not engine.dataInitialized:
        raise RuntimeError("data are not set")
    if not engine.locationsInitialized:
        raise RuntimeError("locations are not set")
    if engine.precision is None:
        raise RuntimeError("precision is not set")


def getLogLikelihood(engine: Engine) -> float:
    """R/mds.R:10-30: 0.5 (log tau - log 2 pi) N (N - 1) / 2 - sumOfIncrements."""
    _check_ready(engine)
    s = engine.engine.getSumOfIncrements()
    n = engine.locationCount
    observationCount = (n * (n - 1)) / 2
    return 0.5 * (math.log(engine.precision) - math.log(2 * math.pi)) * observationCount - s


def getGradient(engine: Engine) -> np.ndarray:
    """R/mds.R:41-57: the N x P gradient (row-major from the engine)."""
    _check_ready(engine)
    g = engine.engine.getLogLikelihoodGradient(engine.locationCount * engine.embeddingDimension)
    return g.reshape(engine.locationCount, engine.embeddingDimension)
