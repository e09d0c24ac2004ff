"""GPU tests of storeState / restoreState / acceptState through the C ABI,
sequence by sequence (src/NewMultiDimensionalScaling.hpp:205-255): after
every restore the engine must hold the stored locations, whatever was
written in between (every location, one row, nothing) and however many
stores and accepts came before.  The stored state stays in its buffer until
something writes part of it, so these sequences reach each way the buffers
can be taken.

Each check forces a full evaluation (makeDirty) at the state the engine
holds and compares it with a separate engine given those locations: the
same kernels on the same data, so the sum and the gradient must be bitwise
equal; the sums are also held to the serial oracle (FP64 1e-12, FP32 2e-6).
"""
import numpy as np
import pytest

from massivemds_amd import datagen
from massivemds_amd.engine import Flags, MultiDimensionalScaling

pytestmark = pytest.mark.gpu

TAU = 1.3


def _flags(fp32):
    return Flags.OPENCL | Flags.LEFT_TRUNCATION | (Flags.FLOAT if fp32 else 0)


class _Case:
    def __init__(self, n, p, fp32, seed):
        self.n, self.p, self.fp32 = n, p, fp32
        self.delta = datagen.euclid10(n, seed, 0.1, 0.05)
        self.e = self._engine()
        self.fresh = self._engine()

    def _engine(self):
        e = MultiDimensionalScaling(self.p, self.n, _flags(self.fp32), 0)
        e.setPairwiseData(self.delta)
        e.setParameters([TAU])
        return e

    def locations(self, seed):
        return datagen.locations(self.n, self.p, seed)

    def expect(self, oracle, x, label):
        """The engine's state, evaluated afresh, is the one at locations x."""
        print(f"store/restore {label} n={self.n} p={self.p} fp32={self.fp32}", flush=True)
        self.e.makeDirty()
        s = self.e.getSumOfIncrements()
        g = self.e.getLogLikelihoodGradient()
        self.fresh.updateLocations(-1, x)
        self.fresh.makeDirty()
        assert s == self.fresh.getSumOfIncrements(), label
        assert np.array_equal(g, self.fresh.getLogLikelihoodGradient()), label
        ref = oracle.sum_of_increments(self.delta, x, TAU, True)
        tol = 2e-6 if self.fp32 else 1e-12
        assert abs(s - ref) <= tol * abs(ref), (label, s, ref)

    def close(self):
        self.e.close()
        self.fresh.close()


@pytest.mark.parametrize("n,p,fp32", [(300, 2, False), (257, 3, False), (300, 2, True)])
def test_restore_returns_the_stored_locations(oracle, n, p, fp32):
    c = _Case(n, p, fp32, 41)
    e = c.e
    try:
        x1, x2, x3 = c.locations(42), c.locations(43), c.locations(44)
        e.updateLocations(-1, x1)
        e.getSumOfIncrements()

        # nothing written between store and restore
        e.storeState()
        e.restoreState()
        c.expect(oracle, x1, "restore after nothing")

        # every location written, then restored
        e.storeState()
        e.updateLocations(-1, x2)
        c.expect(oracle, x2, "after a full upload")
        e.restoreState()
        c.expect(oracle, x1, "restore after a full upload")

        # one row moved (O(N) update of a known sum), then restored
        e.getSumOfIncrements()
        e.storeState()
        row = x1[n // 3] + 0.25
        e.updateLocations(n // 3, row)
        moved = x1.copy()
        moved[n // 3] = row
        s_row = e.getSumOfIncrements()
        ref = oracle.sum_of_increments(c.delta, moved, TAU, True)
        assert abs(s_row - ref) <= (2e-6 if fp32 else 1e-12) * abs(ref)
        e.restoreState()
        c.expect(oracle, x1, "restore after a row update")

        # one row moved but never evaluated: the pending write is given up
        e.storeState()
        e.updateLocations(7, x1[7] - 0.5)
        e.restoreState()
        c.expect(oracle, x1, "restore after a pending row")

        # accepted moves keep the stored state of the last store
        e.storeState()
        e.updateLocations(-1, x2)
        e.getSumOfIncrements()
        e.acceptState()
        e.updateLocations(-1, x3)
        e.restoreState()
        c.expect(oracle, x1, "restore after accept and another upload")

        # two stores in a row, then a full upload
        e.updateLocations(-1, x2)
        e.storeState()
        e.storeState()
        e.updateLocations(-1, x3)
        e.restoreState()
        c.expect(oracle, x2, "restore after two stores")

        # a row written in place after a store, then stored again and replaced
        e.updateLocations(-1, x3)
        e.getSumOfIncrements()
        e.storeState()
        e.updateLocations(5, x3[5] + 1.0)
        e.getSumOfIncrements()
        kept = x3.copy()
        kept[5] = x3[5] + 1.0
        e.storeState()
        e.updateLocations(-1, x1)
        e.getSumOfIncrements()
        e.restoreState()
        c.expect(oracle, kept, "restore to a state stored after a row move")
        e.restoreState()  # the previous store was replaced: the same state again
        c.expect(oracle, kept, "second restore")
    finally:
        c.close()
