"""Portable synthetic inputs for the MDS likelihood.

The reference benchmark draws its data with ``std::mt19937`` and
``std::normal_distribution`` (src/benchmark.cpp:63-81, 174-189), whose output is
implementation-defined.  These generators are instead fully specified so that
the same seed gives the same inputs in any process that uses this module:

* uniforms: a splitmix64 stream per (seed, key); each 64-bit output ``u``
  maps to ``((u >> 11) + 0.5) * 2**-53`` in (0, 1);
* normals: Box-Muller on consecutive uniform pairs (``r = sqrt(-2 log u1)``,
  ``theta = 2 pi u2``, giving ``r cos theta`` then ``r sin theta``).

Observation generators:

* ``euclid10``: rows of Z are i.i.d. N(0, I_10); delta_ij = ||z_i - z_j||_2,
  mirrored, zero diagonal.  Optional additive N(0, sigma^2) noise drawn once
  per i < j and clipped at 0, and a symmetric missing (NaN) fraction.
* ``chisq``: delta_ij = z^2, z ~ N(0, 1) per i < j, as in the reference
  benchmark (src/benchmark.cpp:176-189), with its optional 25% missingness.
"""
from __future__ import annotations

import numpy as np

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

DATA_SEED = 666  # src/benchmark.cpp:63, R/Untitled.R:546
LOCATION_SEED = 667


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _stream_state(seed: int, key: int) -> np.uint64:
    s = np.array([(int(seed) * 0x100000001B3 + int(key) * 0xC2B2AE3D27D4EB4F) & 0xFFFFFFFFFFFFFFFF],
                 dtype=np.uint64)
    return _mix(s)[0]


def uniforms(seed: int, key: int, count: int, offset: int = 0) -> np.ndarray:
    """``count`` uniforms in (0, 1) from stream (seed, key), starting at ``offset``."""
    state = _stream_state(seed, key)
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = _mix(state + idx * _GOLDEN)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


def normals(seed: int, key: int, count: int) -> np.ndarray:
    """``count`` standard normals from stream (seed, key) by Box-Muller."""
    pairs = (count + 1) // 2
    u = uniforms(seed, key, 2 * pairs)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    t = 2.0 * np.pi * u2
    out = np.empty(2 * pairs)
    out[0::2] = r * np.cos(t)
    out[1::2] = r * np.sin(t)
    return out[:count]


def locations(n: int, p: int, seed: int = LOCATION_SEED, key: int = 1) -> np.ndarray:
    """Row-major N x P latent locations drawn from N(0, I_P)."""
    return normals(seed, key, n * p).reshape(n, p)


def _upper_pairs_mask(n: int, frac: float, seed: int, key: int) -> np.ndarray:
    """Symmetric boolean N x N matrix: True for i != j pairs drawn missing."""
    iu = np.triu_indices(n, 1)
    u = uniforms(seed, key, iu[0].size)
    m = np.zeros((n, n), dtype=bool)
    m[iu] = u < frac
    return m | m.T


def _distances(z: np.ndarray, block: int) -> np.ndarray:
    """||z_i - z_j||_2 for all i, j, a block of rows at a time into buffers
    allocated once.  The squared coordinate differences are added in the
    order numpy's pairwise summation uses for a row of ``dim`` values
    (``np.sum(d * d, axis=-1)``: eight partial sums combined as a tree, the
    rest added in turn), so the values are those of that expression bit for
    bit."""
    n, dim = z.shape
    zt = np.ascontiguousarray(z.T)
    delta = np.empty((n, n))
    rows = min(block, n)
    sq = np.empty((dim, rows, n))
    a = np.empty((rows, n))
    b = np.empty((rows, n))
    for r0 in range(0, n, block):
        r1 = min(n, r0 + block)
        m = r1 - r0
        q = sq[:, :m]
        for k in range(dim):
            np.subtract(zt[k, r0:r1, None], zt[k, None, :], out=q[k])
            np.multiply(q[k], q[k], out=q[k])
        out = delta[r0:r1]
        if dim >= 8:
            am, bm = a[:m], b[:m]
            np.add(q[0], q[1], out=am)
            np.add(q[2], q[3], out=bm)
            np.add(am, bm, out=am)
            np.add(q[4], q[5], out=bm)
            np.add(q[6], q[7], out=out)
            np.add(bm, out, out=bm)
            np.add(am, bm, out=out)
            for k in range(8, dim):
                np.add(out, q[k], out=out)
        else:
            np.copyto(out, q[0])
            for k in range(1, dim):
                np.add(out, q[k], out=out)
        np.sqrt(out, out=out)
    return delta


def euclid10(n: int, seed: int = DATA_SEED, noise_sd: float = 0.0, missing: float = 0.0,
             dim: int = 10, block: int = 512) -> np.ndarray:
    """Row-major N x N observations: distances of N(0, I_dim) points.

    With ``noise_sd > 0`` each i < j distance gets N(0, noise_sd^2) noise (drawn
    once, mirrored) and is clipped at 0; ``missing`` is the fraction of i < j
    pairs set to NaN on both sides.
    """
    z = normals(seed, 1, n * dim).reshape(n, dim)
    delta = _distances(z, block)
    np.fill_diagonal(delta, 0.0)
    if noise_sd > 0.0:
        iu = np.triu_indices(n, 1)
        eps = noise_sd * normals(seed, 2, iu[0].size)
        upper = np.maximum(delta[iu] + eps, 0.0)
        delta[iu] = upper
        delta.T[iu] = upper
    if missing > 0.0:
        delta[_upper_pairs_mask(n, missing, seed, 3)] = np.nan
    return delta


def chisq(n: int, seed: int = DATA_SEED, missing: bool = False) -> np.ndarray:
    """Reference-benchmark style observations: z^2 per i < j (src/benchmark.cpp:176-189)."""
    iu = np.triu_indices(n, 1)
    z = normals(seed, 4, iu[0].size)
    delta = np.zeros((n, n))
    delta[iu] = z * z
    delta.T[iu] = z * z
    if missing:
        delta[_upper_pairs_mask(n, 0.25, seed, 5)] = np.nan
    return delta


def batch_specs(count: int = 64, n_lo: int = 500, n_hi: int = 8000, dims=(2, 5, 10), seed: int = 999):
    """(N_k, P_k, tau_k, seed_k) of the uneven batch: N log-uniform in
    [n_lo, n_hi], P uniform on `dims`, tau uniform in [0.5, 4], dataset seeds
    1000 + k."""
    u = uniforms(seed, 1, 3 * count).reshape(count, 3)
    n = np.rint(np.exp(np.log(n_lo) + u[:, 0] * (np.log(n_hi) - np.log(n_lo)))).astype(int)
    p = np.asarray(dims)[np.minimum((u[:, 1] * len(dims)).astype(int), len(dims) - 1)]
    tau = 0.5 + 3.5 * u[:, 2]
    return [(int(n[k]), int(p[k]), float(tau[k]), 1000 + k) for k in range(count)]


def batch_dataset(n: int, p: int, seed: int):
    """Observations and locations of one batch dataset (euclid10, N(0, I_P))."""
    return euclid10(n, seed=seed), locations(n, p, seed=seed, key=2)


def config(name: str):
    """Inputs of the named workload: (delta, X, precision, truncation)."""
    if name == "n1000_p2":  # reference CPU-runnable case
        n, p = 1000, 2
        return euclid10(n), locations(n, p), 1.0, True
    if name == "n10000_p2":  # scale + precision case
        n, p = 10000, 2
        return euclid10(n, noise_sd=0.1, missing=0.05), locations(n, p), 2.0, True
    raise KeyError(name)
