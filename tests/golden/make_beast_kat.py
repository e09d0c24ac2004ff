"""Writes tests/golden/beast_h3_kat.npz: the external known answer for the
non-truncated MDS log-likelihood, taken from the data files the reference
ships (inst/extdata).

* observations: inst/extdata/h3_Deff.txt, a 372 x 372 tab-separated distance
  matrix with 4-decimal entries (stored here as integers x 1e4, exact);
* locations: the `loc` annotations of the 372 tips of the BEAST state in
  inst/extdata/large.trees (tree line of STATE_520000), mapped to the matrix
  rows through the NEXUS Translate block, the label matching that
  R/Untitled.R:358-374 (getdata) performs;
* expected value: the logged `mdsLikelihood` and `mds.residual.precision` of
  the same state in inst/extdata/large.log (last line), and likewise for the
  second state in inst/extdata/h3_large_geo_BMDS_Country_HMC.{trees,log}
  (whose precision is logged to 7 digits only, so it pins less).

Run from the repository root in the build container (the reference tree is
read as data; nothing of it is imported or executed):
    python tests/golden/make_beast_kat.py
"""
import re
from pathlib import Path

import numpy as np

EXT = Path("/root/reference/inst/extdata")
OUT = Path(__file__).resolve().parent / "beast_h3_kat.npz"


def read_matrix(path):
    lines = path.read_text().splitlines()
    labels, rows = [], []
    for line in lines[1:]:
        if not line.strip():
            continue
        parts = line.split("\t")
        labels.append(parts[0])
        rows.append([int(round(float(v) * 10000)) for v in parts[1:]])
    m = np.array(rows, dtype=np.int32)
    assert m.shape == (len(labels), len(labels))
    # the integer form is exact: every entry has 4 decimals
    for line in lines[1:4]:
        for v in line.split("\t")[1:]:
            assert re.fullmatch(r"\s*\d+\.\d{4}", v), v
    return labels, m


def read_tips(path):
    text = path.read_text()
    block = text[text.index("Translate"):]
    block = block[: block.index(";")]
    translate = {}
    for num, label in re.findall(r"^\s*(\d+)\s+([^,\s]+),?\s*$", block, flags=re.M):
        translate[int(num)] = label
    tree = [l for l in text.splitlines() if l.strip().startswith("tree STATE_")][-1]
    tips = {}
    for num, a, b in re.findall(r"[(,](\d+)\[&loc=\{([^,}]+),([^}]+)\}\]", tree):
        tips[translate[int(num)]] = (float(a), float(b))
    return tips


def read_log(path):
    rows = [l.split("\t") for l in path.read_text().splitlines() if l and not l.startswith("#")]
    header, last = rows[0], rows[-1]
    return (float(last[header.index("mdsLikelihood")]),
            float(last[header.index("mds.residual.precision")]))


def one(trees, log, labels):
    tips = read_tips(EXT / trees)
    assert set(tips) == set(labels), "tree tips and matrix rows differ"
    x = np.array([tips[l] for l in labels])
    loglik, tau = read_log(EXT / log)
    return x, loglik, tau


def main():
    labels, d = read_matrix(EXT / "h3_Deff.txt")
    x1, l1, t1 = one("large.trees", "large.log", labels)
    x2, l2, t2 = one("h3_large_geo_BMDS_Country_HMC.trees", "h3_large_geo_BMDS_Country_HMC.log", labels)
    np.savez_compressed(OUT, delta_x1e4=d, locations=x1, precision=t1, loglik=l1,
                        locations2=x2, precision2=t2, loglik2=l2)
    print(OUT, d.shape, l1, t1, l2, t2)


if __name__ == "__main__":
    main()
