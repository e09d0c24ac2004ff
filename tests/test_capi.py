"""CPU checks of the drop-in boundary: libmds_hip.so loads, exports every
function include/mds_hip.h declares, and fails cleanly (status + message, no
crash, no CPU fallback) where no GPU is present."""
import ctypes
import os
import re
import subprocess
import sys
from pathlib import Path

import pytest

from massivemds_amd import engine as eng
from massivemds_amd import mds

ROOT = Path(__file__).resolve().parents[1]


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mds_hip_\w+)\s*\(", text)))


def _lib():
    if not eng.library_path().exists():
        from massivemds_amd import build

        build.build()
    return ctypes.CDLL(str(eng.library_path()))


@pytest.mark.parametrize("header,minimum", [("mds_hip.h", 20), ("mds_hip_singleton.h", 14)])
def test_header_symbols_exported(header, minimum):
    lib = _lib()
    names = _declared(header)
    assert len(names) >= minimum
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_flags_match_reference_bits():
    # src/flags.h:6-14
    assert eng.Flags.FLOAT == 4 and eng.Flags.TBB == 8 and eng.Flags.OPENCL == 16
    assert eng.Flags.LEFT_TRUNCATION == 32 and eng.Flags.SSE == 128
    assert eng.Flags.AVX == 256 and eng.Flags.AVX512 == 512
    text = (ROOT / "include" / "mds_hip.h").read_text()
    for name, bit in [("FLOAT", 2), ("TBB", 3), ("OPENCL", 4), ("LEFT_TRUNCATION", 5), ("SSE", 7),
                      ("AVX", 8), ("AVX512", 9), ("HIP", 10)]:
        assert re.search(rf"MDS_HIP_FLAG_{name} \(1L << {bit}\)", text), name


def test_max_dimension():
    lib = eng.load_library()
    assert lib.mds_hip_max_dimension() == 16


def test_create_without_gpu_reports_error():
    lib = eng.load_library()
    if eng.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(eng.MdsError) as info:
        eng.MultiDimensionalScaling(2, 10, eng.Flags.OPENCL)
    assert info.value.status == 2
    assert "device" in str(info.value).lower()


def test_invalid_arguments_rejected_before_device():
    for p in (0, 17):
        with pytest.raises(eng.MdsError) as info:
            eng.MultiDimensionalScaling(p, 10, eng.Flags.OPENCL)
        assert info.value.status == 1
    with pytest.raises(eng.MdsError):
        eng.MultiDimensionalScaling(2, 0, eng.Flags.OPENCL)


def test_null_handle_calls_fail():
    lib = eng.load_library()
    assert lib.mds_hip_store_state(None) == 1
    assert lib.mds_hip_destroy(None) == 1
    assert b"null" in lib.mds_hip_last_error(None)


def test_cpu_routes_are_not_served():
    with pytest.raises(ValueError):
        eng.factory(2, 10, eng.Flags.TBB, 0, 4)
    with pytest.raises(ValueError):
        mds.createEngine(2, 10, truncation=True, gpu=0)


def test_singleton_registry_errors_without_gpu():
    """The BEAST bridge's registry refuses bad instance numbers and CPU flag
    sets with a status and a message (no GPU needed)."""
    lib = _lib()
    lib.mds_hip_singleton_last_error.restype = ctypes.c_char_p
    lib.mds_hip_singleton_get_sum_of_increments.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    s = ctypes.c_double()
    assert lib.mds_hip_singleton_get_sum_of_increments(12345, ctypes.byref(s)) == 1
    assert b"no instance" in lib.mds_hip_singleton_last_error(-1)
    inst = ctypes.c_int(-7)
    lib.mds_hip_singleton_initialize.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_int,
                                                 ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    assert lib.mds_hip_singleton_initialize(2, 10, 1 << 5, 0, 1, ctypes.byref(inst)) == 1  # no GPU bit
    assert b"GPU engine only" in lib.mds_hip_singleton_last_error(-1)
    assert inst.value == -7


def test_singleton_harness_builds(tmp_path):
    from tests.singleton_harness import harness

    exe = harness(tmp_path)
    assert exe.exists()


def test_missing_library_fails_loudly(tmp_path):
    """Without the built HIP library the engine raises at construction: there
    is no CPU fallback to run instead."""
    code = ("import numpy as np\n"
            "from massivemds_amd.engine import Flags, MultiDimensionalScaling\n"
            "try:\n"
            "    MultiDimensionalScaling(2, 10, Flags.OPENCL, 0)\n"
            "except RuntimeError as e:\n"
            "    print('raised:', e)\n"
            "else:\n"
            "    print('no error')\n")
    env = dict(os.environ, MDS_HIP_LIBRARY=str(tmp_path / "absent" / "libmds_hip.so"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env,
                       cwd=Path(__file__).resolve().parents[1])
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("raised:") and "is missing" in r.stdout
