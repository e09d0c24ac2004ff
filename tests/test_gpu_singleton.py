"""GPU test of the BEAST bridge (include/mds_hip_singleton.h) through a C
program that makes BEAST's call sequence (tests/c/beast_singleton_harness.c):
results against the CPU oracle, store/restore/accept, error statuses."""
import subprocess

import numpy as np
import pytest

from massivemds_amd import datagen
from tests.singleton_harness import harness

pytestmark = pytest.mark.gpu


def _read(path):
    out = {}
    for line in path.read_text().splitlines():
        key, *vals = line.split()
        out[key] = [float(v) for v in vals]
    return out


def test_beast_call_sequence(tmp_path, oracle):
    n, p, tau = 150, 3, 1.3
    delta = datagen.euclid10(n, 51, 0.1, 0.05)
    x = datagen.locations(n, p, 52)
    delta.tofile(tmp_path / "delta.bin")
    x.tofile(tmp_path / "x.bin")
    exe = harness(tmp_path)
    print("running the BEAST call-sequence harness", flush=True)
    proc = subprocess.run([str(exe), str(tmp_path / "delta.bin"), str(tmp_path / "x.bin"), str(n), str(p),
                           repr(tau), str(tmp_path / "out.txt")], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr
    r = _read(tmp_path / "out.txt")
    assert r["instances"] == [0.0, 0.0, 1.0]
    moved = x.copy()
    moved[3] += 0.1 * np.arange(1, p + 1)
    s0 = oracle.sum_of_increments(delta, x, tau, True)
    s1 = oracle.sum_of_increments(delta, moved, tau, True)
    g0 = oracle.gradient(delta, x, tau, True)
    g1 = oracle.gradient(delta, moved, tau, True)
    assert abs(r["sum0"][0] - s0) <= 1e-12 * abs(s0)
    assert np.max(np.abs(np.array(r["grad0"]) - g0.ravel())) <= 1e-12 * np.max(np.abs(g0))
    assert abs(r["sum_moved"][0] - s1) <= 1e-12 * abs(s1)
    assert r["sum_restored"][0] == r["sum0"][0]
    assert r["sum_accepted"][0] == r["sum_moved"][0]
    assert np.max(np.abs(np.array(r["grad_accepted"]) - g1.ravel())) <= 1e-12 * np.max(np.abs(g1))
    assert r["internal_dimension"] == [float(p)]
    assert abs(r["sum_dirty"][0] - s1) <= 1e-12 * abs(s1)
    assert r["bad_instance"] == [1.0] and r["bad_length"] == [1.0] and r["cpu_flags"] == [1.0]
    assert r["after_release"] == [1.0]
