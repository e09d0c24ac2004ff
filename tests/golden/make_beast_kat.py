"""Writes tests/golden/beast_h3_kat.npz: the external known answer for the
non-truncated MDS log-likelihood, taken from the data files the reference
ships (inst/extdata).

* observations: inst/extdata/h3_Deff.txt, a 372 x 372 tab-separated distance
  matrix with 4-decimal entries (stored here as integers x 1e4, exact);
* locations: the `loc` annotations of the 372 tips of the BEAST state in
  inst/extdata/large.trees (tree line of STATE_520000), mapped to the matrix
  rows through the NEXUS Translate block, the label matching that
  R/Untitled.R:358-374 (getdata) performs (read with massivemds_amd.beastio);
* expected value: the logged `mdsLikelihood` and `mds.residual.precision` of
  the same state in inst/extdata/large.log (last line), and likewise for the
  second state in inst/extdata/h3_large_geo_BMDS_Country_HMC.{trees,log}
  (whose precision is logged to 7 digits only, so it pins less).

Run in the build container (the reference tree is read as data; nothing of
it is imported or executed):
    python tests/golden/make_beast_kat.py
"""
import re
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))  # the repository root
from massivemds_amd import beastio  # noqa: E402

EXT = Path("/root/reference/inst/extdata")
OUT = Path(__file__).resolve().parent / "beast_h3_kat.npz"


def read_matrix(path):
    labels, d = beastio.read_distance_matrix(path)
    m = np.round(d * 10000).astype(np.int32)
    # the integer form is exact: every entry has 4 decimals
    for line in path.read_text().splitlines()[1:4]:
        for v in line.split("\t")[1:]:
            assert re.fullmatch(r"\s*\d+\.\d{4}", v), v
    assert np.array_equal(m.astype(np.float64) / 10000.0, d)
    return labels, m


def one(trees, log, labels):
    state = beastio.load_state(EXT / "h3_Deff.txt", EXT / trees)
    x = state.locations[state.inverse_permutation]  # tips in the matrix's row order
    assert [state.labels[k] for k in state.inverse_permutation] == labels
    cols = beastio.read_log(EXT / log)
    return x, float(cols["mdsLikelihood"][-1]), float(cols["mds.residual.precision"][-1])


def main():
    labels, d = read_matrix(EXT / "h3_Deff.txt")
    x1, l1, t1 = one("large.trees", "large.log", labels)
    x2, l2, t2 = one("h3_large_geo_BMDS_Country_HMC.trees", "h3_large_geo_BMDS_Country_HMC.log", labels)
    np.savez_compressed(OUT, delta_x1e4=d, locations=x1, precision=t1, loglik=l1,
                        locations2=x2, precision2=t2, loglik2=l2)
    print(OUT, d.shape, l1, t1, l2, t2)


if __name__ == "__main__":
    main()
