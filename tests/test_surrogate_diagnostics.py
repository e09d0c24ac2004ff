"""CPU tests of the surrogate diagnostics in massivemds_amd/surrogate.py: the
analytic Hessian is pinned by central differences of the oracle's gradient,
the per-location ceiling by a case solved by hand."""
import numpy as np

from massivemds_amd import datagen
from massivemds_amd.surrogate import per_location_ceiling, residual_hessian


def test_residual_hessian_matches_oracle_gradient_differences(oracle):
    n, p, tau = 7, 3, 1.7
    delta = datagen.euclid10(n, 5, 0.1, 0.0)
    delta[1, 4] = delta[4, 1] = np.nan       # a missing pair
    delta[2, 5] = np.nan                     # observed in one direction only
    x = datagen.normals(6, 1, n * p).reshape(n, p)
    h = residual_hessian(delta, x, tau)
    fd = np.empty_like(h)
    eps = 1e-6
    for k in range(n * p):
        xp, xm = x.copy().ravel(), x.copy().ravel()
        xp[k] += eps
        xm[k] -= eps
        gp = oracle.gradient(delta, xp.reshape(n, p), tau, False).ravel()
        gm = oracle.gradient(delta, xm.reshape(n, p), tau, False).ravel()
        fd[:, k] = -(gp - gm) / (2 * eps)
    assert np.max(np.abs(h - fd)) <= 1e-6 * np.max(np.abs(h))
    # symmetric observations: the Hessian of minus the log-likelihood
    delta[5, 2] = np.nan
    hs = residual_hessian(delta, x, tau)
    assert np.allclose(hs, hs.T, rtol=0, atol=1e-12)
    assert not np.allclose(h, h.T)


def test_ceiling_two_locations_by_hand():
    # H = [[a, b], [b, a]] (P = 1): Sigma_11 = a / (a^2 - b^2), so the
    # fraction is (a^2 - b^2) / a^2
    a, b = 3.0, 1.2
    assert abs(per_location_ceiling(np.array([[a, b], [b, a]]), 1) - (1 - b * b / (a * a))) < 1e-14


def test_ceiling_block_diagonal_is_one():
    rng = np.random.default_rng(0)
    blocks = []
    for _ in range(4):
        m = rng.normal(size=(3, 3))
        blocks.append(m @ m.T + 3 * np.eye(3))
    h = np.zeros((12, 12))
    for i, blk in enumerate(blocks):
        h[3 * i:3 * i + 3, 3 * i:3 * i + 3] = blk
    assert abs(per_location_ceiling(h, 3) - 1.0) < 1e-12
    h[0, 5] = h[5, 0] = 1.0
    assert per_location_ceiling(h, 3) < 1.0


def test_ceiling_leaves_out_flat_directions():
    # H = [[1, -1], [-1, 1]] (P = 1) is flat along (1, 1): on its range
    # x_1 = -x_2, so x_i alone fixes gradient row i and the fraction is 1
    assert abs(per_location_ceiling(np.array([[1.0, -1.0], [-1.0, 1.0]]), 1) - 1.0) < 1e-12
    # adding a flat direction to a coupled H leaves the fraction unchanged
    a, b = 3.0, 1.2
    h = np.zeros((3, 3))
    h[:2, :2] = [[a, b], [b, a]]
    assert abs(per_location_ceiling(h, 1) - (1 - b * b / (a * a)) * 2 * a / (2 * a)) < 1e-12


class _OracleEngine:
    """The engine methods map_locations calls, on the CPU oracle (tests only)."""

    def __init__(self, oracle, delta, precision, truncated):
        self.o, self.delta, self.tau, self.trunc = oracle, delta, precision, truncated

    def setParameters(self, v):
        self.tau = v[0]

    def updateLocations(self, index, x):
        self.x = np.asarray(x, dtype=np.float64).reshape(self.delta.shape[0], -1)

    def getSumOfIncrements(self):
        return self.o.sum_of_increments(self.delta, self.x, self.tau, self.trunc)

    def getLogLikelihoodGradient(self):
        return self.o.gradient(self.delta, self.x, self.tau, self.trunc).ravel()


def _mode_and_hessian(oracle, n, p):
    from massivemds_amd.surrogate import map_locations

    delta = datagen.euclid10(n, missing=0.02)
    eng = _OracleEngine(oracle, delta, 1.0, False)
    x = map_locations(eng, datagen.normals(667, 3, n * p).reshape(n, p))
    obs = ~np.isnan(delta)
    np.fill_diagonal(obs, False)
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    tau = 0.5 * obs.sum() / (0.5 * np.sum((delta - d)[obs] ** 2))  # pairs / sum of squared residuals
    eng.setParameters([tau])
    x = map_locations(eng, x)
    g = eng.getLogLikelihoodGradient().reshape(n, p)
    assert np.linalg.norm(g - x) <= 1e-3 * np.linalg.norm(x)  # a mode
    return residual_hessian(delta, x, tau) + np.eye(n * p), tau


def test_per_location_ceiling_of_the_surrogate_workload(oracle):
    """The surrogate workload's posterior family (distances of N(0, I_10)
    points, 2% missing, plug-in precision, at its mode): a per-location model
    of the gradient rows can explain well under half of the gradient's
    variance at P = 5, and its stiffest direction is collective, with
    curvature N tau (all distances scaling together)."""
    n = 500
    h5, tau5 = _mode_and_hessian(oracle, n, 5)
    c5 = per_location_ceiling(h5, 5)
    top = np.linalg.eigvalsh(0.5 * (h5 + h5.T))[-1]
    h2, _ = _mode_and_hessian(oracle, n, 2)
    c2 = per_location_ceiling(h2, 2)
    print(f"ceiling P=5 {c5:.3f}, P=2 {c2:.3f}; largest curvature {top:.1f} vs N tau {n * tau5:.1f}")
    assert c5 < 0.5
    assert c2 > 0.75
    assert abs(top / (n * tau5) - 1.0) < 0.05
