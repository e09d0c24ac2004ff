"""Reading BEAST output and distance data for the MDS likelihood: the data
ingestion of the reference's R layer, for real-data runs of the engine.

* ``read_distance_matrix``: a labelled, tab-separated N x N matrix with a
  header line and the row label in the first column, as
  ``read.table(sep = "\\t", header = TRUE, row.names = 1)`` reads it in
  ``getdata`` (R/Untitled.R:358-361).
* ``read_nexus_trees``: the TREES block of a NEXUS file with its Translate
  table (tip numbers to labels, quotes removed) and every tree with its
  name, its state annotations (``[&lnP=..., precision={...}]``) and its Newick
  string, as ``read.annotated.nexus`` splits them (R/AnnotatedTreeReader.R:351-420).
* ``tip_annotations``: one annotated trait of every tip (``[&loc={x,y}]``),
  keyed by label, as ``readbeast`` takes the tip locations
  (R/Untitled.R:326-333).
* ``read_log``: the columns of a BEAST log (tab-separated, ``#`` comments).
* ``load_state``: observations, tip locations and the tree's annotations of
  one state, with the observations permuted into the tips' order and the
  inverse permutation kept, as ``getdata`` does (R/Untitled.R:358-374).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_NUMBER = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def read_distance_matrix(path) -> tuple[list[str], np.ndarray]:
    lines = [l for l in Path(path).read_text().splitlines() if l.strip()]
    labels, rows = [], []
    for line in lines[1:]:  # the first line holds the column labels
        parts = line.split("\t")
        labels.append(parts[0].strip().strip("\"'"))
        rows.append([float(v) if v.strip() not in ("NA", "NaN", "") else np.nan for v in parts[1:]])
    m = np.array(rows, dtype=np.float64)
    if m.shape != (len(labels), len(labels)):
        raise ValueError(f"{path}: expected a square labelled matrix, read {m.shape} for {len(labels)} labels")
    return labels, m


def _parse_annotations(text: str) -> dict:
    """``&key=value,key={v1,v2,...}`` (BEAST comment syntax) to a dict of
    floats, float lists or strings."""
    out = {}
    body = text.strip()
    if body.startswith("&"):
        body = body[1:]
    for key, vec, scalar in re.findall(r"([\w.]+)=(?:\{([^}]*)\}|([^,}\]]*))", body):
        if vec:
            out[key] = [float(v) for v in vec.split(",")]
        else:
            try:
                out[key] = float(scalar)
            except ValueError:
                out[key] = scalar.strip("\"'")
    return out


@dataclass
class NexusTree:
    name: str
    annotations: dict
    newick: str


def read_nexus_trees(path) -> tuple[dict[str, str], list[NexusTree]]:
    text = Path(path).read_text()
    lower = text.lower()
    start = lower.index("begin trees;")
    block = text[start:]
    translate = {}
    m = re.search(r"\btranslate\b(.*?);", block, flags=re.I | re.S)
    if m:
        for entry in m.group(1).split(","):
            parts = entry.split()
            if len(parts) >= 2:
                translate[parts[0]] = parts[1].strip("\"'")
    trees = []
    for tm in re.finditer(r"^\s*tree\s+(\S+)\s*(\[[^\]]*\])?\s*=\s*(\[&[^\]]*\])?\s*(.*?);\s*$", block,
                          flags=re.I | re.M):
        name, ann, _, newick = tm.groups()
        trees.append(NexusTree(name, _parse_annotations(ann[1:-1]) if ann else {}, newick.strip()))
    if not trees:
        raise ValueError(f"{path}: no trees in the TREES block")
    return translate, trees


def tip_annotations(tree: NexusTree, translate: dict[str, str], trait: str = "loc") -> dict[str, np.ndarray]:
    """Annotated `trait` of every tip (a node name right after '(' or ',')."""
    tips = {}
    pattern = r"[(,]([^(),:\[\]]+)\[&([^\]]*)\]"
    for tip, ann in re.findall(pattern, tree.newick):
        values = _parse_annotations(ann)
        if trait in values:
            label = translate.get(tip.strip(), tip.strip().strip("\"'"))
            v = values[trait]
            tips[label] = np.atleast_1d(np.asarray(v, dtype=np.float64))
    return tips


def read_log(path) -> dict[str, np.ndarray]:
    rows = [l.split("\t") for l in Path(path).read_text().splitlines() if l.strip() and not l.startswith("#")]
    header = [h.strip() for h in rows[0]]
    cols = {h: [] for h in header}
    for r in rows[1:]:
        for h, v in zip(header, r):
            cols[h].append(v)
    out = {}
    for h, v in cols.items():
        try:
            out[h] = np.array([float(x) for x in v])
        except ValueError:
            out[h] = np.array(v)
    return out


@dataclass
class BeastState:
    labels: list[str]                # tips, in the tree file's order of appearance
    delta: np.ndarray                # N x N observations in the tips' order
    locations: np.ndarray            # N x P tip locations
    annotations: dict                # the tree's state annotations (lnP, precision, ...)
    inverse_permutation: np.ndarray  # matrix row of tip k -> k: delta_file = delta[inv][:, inv]
    tree: str = field(default="")


def load_state(matrix_path, trees_path, state: str | None = None, trait: str = "loc") -> BeastState:
    """One state of a BEAST run: the last tree, or the tree named `state`."""
    labels_file, d = read_distance_matrix(matrix_path)
    translate, trees = read_nexus_trees(trees_path)
    tree = trees[-1] if state is None else next(t for t in trees if t.name == state)
    tips = tip_annotations(tree, translate, trait)
    if set(tips) != set(labels_file):
        missing = sorted(set(labels_file) ^ set(tips))[:5]
        raise ValueError(f"tree tips and matrix rows differ (e.g. {missing})")
    order = list(tips)
    index = {l: k for k, l in enumerate(labels_file)}
    perm = np.array([index[l] for l in order])  # tip k <- matrix row perm[k]
    inverse = np.empty_like(perm)
    inverse[perm] = np.arange(len(perm))
    x = np.array([tips[l] for l in order])
    return BeastState(order, d[np.ix_(perm, perm)], x, tree.annotations, inverse, tree.name)
