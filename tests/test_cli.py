"""The benchmark CLI (massivemds_amd/bin/mds_benchmark), counterpart of the
reference's src/benchmark.cpp, driving the HIP engine through mds::factory
and HipMultiDimensionalScaling (include/mds_hip_engine.hpp)."""
import re
import subprocess

import numpy as np
import pytest

from massivemds_amd import build as hipbuild
from massivemds_amd import datagen

EXE = hipbuild.BIN / "mds_benchmark"


def _exe():
    if not EXE.exists():
        hipbuild.build()
    return str(EXE)


def _run(*args, cwd=None):
    return subprocess.run([_exe(), *args], capture_output=True, text=True, timeout=300, cwd=cwd)


def test_usage():
    r = _run("--help")
    assert r.returncode == 1 and "--truncation" in r.stdout and "--generator" in r.stdout
    r = _run("--bogus")
    assert r.returncode == 1 and "Allowed options" in r.stdout


def test_cpu_flags_are_rejected():
    r = _run("--locations", "10", "--tbb", "2")
    assert r.returncode == 1
    assert "GPU engine only" in r.stderr


def test_conflicting_simd():
    r = _run("--gpu", "1", "--sse", "--avx")
    assert r.returncode == 1 and "more than one SIMD" in r.stderr


def _expected(oracle, n, p, iterations, truncated, tau=1.0, seed=666):
    """The CLI's loop replayed on the oracle with the same portable streams."""
    delta = datagen.chisq(n, seed=seed)
    x = datagen.normals(seed + 1, 1, n * p).reshape(n, p)
    restore = datagen.uniforms(seed, 6, max(iterations, 1))
    loglik = oracle.sum_of_increments(delta, x, tau, truncated)
    sum_grad = oracle.gradient(delta, x, tau, truncated).ravel()[1]
    current = x
    for it in range(iterations):
        proposal = datagen.normals(seed + 1, 2 + it, n * p).reshape(n, p)
        loglik += oracle.sum_of_increments(delta, proposal, tau, truncated)
        if not restore[it] < 0.75:
            current = proposal
        sum_grad += oracle.gradient(delta, current, tau, truncated).ravel()[1]
    return loglik / (iterations + 1), sum_grad / (iterations + 1)


@pytest.mark.gpu
@pytest.mark.parametrize("truncated", [True, False])
def test_cli_matches_oracle(oracle, tmp_path, truncated):
    n, p, iterations = 400, 2, 6
    args = ["--gpu", "1", "--locations", str(n), "--iterations", str(iterations)]
    if truncated:
        args.append("--truncation")
    r = _run(*args, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    ll = float(re.search(r"AvgLogLik = (\S+)", r.stdout).group(1))
    sg = float(re.search(r"AvgSumGradient = (\S+)", r.stdout).group(1))
    e_ll, e_sg = _expected(oracle, n, p, iterations, truncated)
    assert abs(ll - e_ll) <= 1e-11 * abs(e_ll), (ll, e_ll)
    assert abs(sg - e_sg) <= 1e-9 * max(abs(e_sg), 1.0), (sg, e_sg)
    assert (tmp_path / "report.txt").read_text().startswith("0 0 no simd 400 2 6 ")


@pytest.mark.gpu
def test_cli_float_and_internal(oracle, tmp_path):
    r = _run("--gpu", "1", "--float", "--truncation", "--internal", "--locations", "300",
             "--dimension", "5", "--iterations", "2", "--generator", "euclid10", cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "single-precision" in r.stdout
    assert np.isfinite(float(re.search(r"AvgLogLik = (\S+)", r.stdout).group(1)))
