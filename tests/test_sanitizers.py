"""AddressSanitizer and UndefinedBehaviorSanitizer over the CPU code that
the GPU results are checked against and that prepares the kernels' input:
the oracle (serial, row-sample, FP32 and SIMD + OpenMP forms) and the host
tile list and packing (massivemds_amd/csrc/mds_host.hpp).  The drivers in
tests/c/ assert invariants and layout definitions besides running clean."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _run(cmd, env=None, timeout=300):
    proc = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout)
    assert proc.returncode == 0, f"{' '.join(map(str, cmd))}\n{proc.stdout}\n{proc.stderr}"
    return proc


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")
def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "san_oracle"
    _run(["gcc", "-std=c11", *SAN, "-fopenmp", str(ROOT / "tests/c/sanitize_oracle.c"),
          str(ROOT / "oracle/mds_oracle.c"), str(ROOT / "oracle/mds_oracle_simd.c"), "-lmvec", "-lm", "-o", str(exe)])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", OMP_NUM_THREADS="2")
    _run([str(exe)], env=env)


@pytest.mark.skipif(shutil.which("g++") is None or not Path("/opt/rocm/include/hip").exists(),
                    reason="g++ or the HIP headers absent")
def test_host_packing_under_sanitizers(tmp_path):
    exe = tmp_path / "san_host"
    _run(["g++", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{ROOT / 'include'}",
          f"-I{ROOT / 'massivemds_amd/csrc'}", str(ROOT / "tests/c/sanitize_host.cpp"), "-L/opt/rocm/lib",
          "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)])
    # the HIP runtime library is linked (error strings) but not initialised;
    # its allocations at load time are not this code's leaks
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    _run([str(exe)], env=env)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_interface_observation_count_under_ubsan(tmp_path):
    """The restated interface's observation count at N above 46,341, where
    the reference's int product N (N - 1) overflows."""
    exe = tmp_path / "san_interface"
    _run(["g++", "-std=c++17", *SAN, f"-I{ROOT / 'include'}", str(ROOT / "tests/c/interface_ubsan.cpp"),
          "-o", str(exe)])
    _run([str(exe)])


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_tune_cache_round_trip(tmp_path):
    """The autotuner's cache after a forced re-tune: the new record is what a
    later load finds, one line per signature (tests/c/tune_cache_check.cpp)."""
    exe = tmp_path / "tune_cache"
    _run(["g++", "-std=c++17", *SAN, f"-I{ROOT / 'massivemds_amd/csrc'}", str(ROOT / "tests/c/tune_cache_check.cpp"),
          "-o", str(exe)])
    _run([str(exe), str(tmp_path / "tune.cache")])


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_singleton_registry_under_tsan(tmp_path):
    """The BEAST bridge's registry (host code of mds_singleton.hip, built as
    C++ over a stub engine) under ThreadSanitizer: concurrent calls on
    shared instances and a concurrent release_all."""
    exe = tmp_path / "singleton_tsan"
    _run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-fno-sanitize-recover=all",
          f"-I{ROOT / 'include'}", "-x", "c++", str(ROOT / "massivemds_amd/csrc/mds_singleton.hip"), "-x", "none",
          str(ROOT / "tests/c/singleton_tsan.cpp"), "-lpthread", "-o", str(exe)])
    _run([str(exe)], env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
