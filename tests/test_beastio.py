"""CPU tests of the BEAST data ingestion (massivemds_amd/beastio.py) on small
files written here in the formats of the reference's inst/extdata (labelled
tab-separated matrix, NEXUS trees with a Translate block and annotated tips,
tab-separated log); the real 372-location files are covered by the fixture
tests/golden/beast_h3_kat.npz, which make_beast_kat.py writes with this
module."""
import numpy as np
import pytest

from massivemds_amd import beastio

MATRIX = "\t".join(["", "B", "A", "C"]) + "\n" + "\n".join([
    "B\t0.0000\t1.2500\t2.0000",
    "A\t1.2500\t0.0000\tNA",
    "C\t2.0000\t3.5000\t0.0000",
]) + "\n"

TREES = """#NEXUS
Begin taxa;
End;
Begin trees;
\tTranslate
\t\t1 'A',
\t\t2 B,
\t\t3 'C'
\t\t;
tree STATE_0 [&lnP=-10.5,precision={2.0,0.5,3.0}] = [&R] ((1[&loc={0.1,0.2}]:1.0,2[&loc={1.5,-0.5}]:1.0)[&loc={0.0,0.0}]:0.5,3[&loc={-2.0,1.0}]:1.5)[&loc={9.0,9.0}];
tree STATE_100 [&lnP=-9.25,precision={2.5,0.25,3.5}] = [&R] ((1[&loc={0.3,0.4}]:1.0,2[&loc={1.25,-0.75}]:1.0)[&loc={0.0,0.0}]:0.5,3[&loc={-2.5,1.5}]:1.5)[&loc={9.0,9.0}];
End;
"""

LOG = "# BEAST log\n# comment\nstate\tmdsLikelihood\tmds.residual.precision\n0\t-12.5\t1.25\n100\t-11.0\t1.5\n"


@pytest.fixture
def files(tmp_path):
    (tmp_path / "d.txt").write_text(MATRIX)
    (tmp_path / "t.trees").write_text(TREES)
    (tmp_path / "r.log").write_text(LOG)
    return tmp_path


def test_matrix(files):
    labels, d = beastio.read_distance_matrix(files / "d.txt")
    assert labels == ["B", "A", "C"]
    assert d[0, 1] == 1.25 and np.isnan(d[1, 2]) and d[2, 1] == 3.5


def test_trees_and_tips(files):
    translate, trees = beastio.read_nexus_trees(files / "t.trees")
    assert translate == {"1": "A", "2": "B", "3": "C"}
    assert [t.name for t in trees] == ["STATE_0", "STATE_100"]
    assert trees[1].annotations == {"lnP": -9.25, "precision": [2.5, 0.25, 3.5]}
    tips = beastio.tip_annotations(trees[0], translate)
    assert set(tips) == {"A", "B", "C"}  # internal nodes and the root are not tips
    assert np.array_equal(tips["B"], [1.5, -0.5]) and np.array_equal(tips["C"], [-2.0, 1.0])


def test_state_permutation(files):
    s = beastio.load_state(files / "d.txt", files / "t.trees")  # the last tree
    assert s.tree == "STATE_100" and s.labels == ["A", "B", "C"]
    assert np.array_equal(s.locations, [[0.3, 0.4], [1.25, -0.75], [-2.5, 1.5]])
    # observations in the tips' order (getdata, R/Untitled.R:358-374)
    assert s.delta[0, 1] == 1.25 and s.delta[1, 0] == 1.25  # A-B
    assert s.delta[2, 0] == 3.5 and np.isnan(s.delta[0, 2])  # C-A, A-C
    _, d = beastio.read_distance_matrix(files / "d.txt")
    inv = s.inverse_permutation
    assert np.array_equal(s.delta[np.ix_(inv, inv)], d, equal_nan=True)
    s0 = beastio.load_state(files / "d.txt", files / "t.trees", state="STATE_0")
    assert s0.annotations["lnP"] == -10.5


def test_log(files):
    cols = beastio.read_log(files / "r.log")
    assert cols["mdsLikelihood"][-1] == -11.0 and cols["mds.residual.precision"][0] == 1.25


def test_mismatched_labels(files, tmp_path):
    (tmp_path / "d2.txt").write_text(MATRIX.replace("C\t", "D\t"))
    with pytest.raises(ValueError):
        beastio.load_state(tmp_path / "d2.txt", files / "t.trees")
