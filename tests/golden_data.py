"""Loaders for the committed golden fixtures under tests/golden/."""
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def beast_kat():
    """BEAST known answer (see tests/golden/make_beast_kat.py)."""
    z = np.load(GOLDEN / "beast_h3_kat.npz")
    delta = z["delta_x1e4"].astype(np.float64) / 10000.0
    return {
        "delta": delta,
        "x": z["locations"],
        "precision": float(z["precision"]),
        "loglik": float(z["loglik"]),
        "x2": z["locations2"],
        "precision2": float(z["precision2"]),
        "loglik2": float(z["loglik2"]),
    }
