"""bench.py's command line on the CPU: more than one GPU is refused with a
message before any workload (or the GPU) is touched."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("extra", [[], ["--workload", "batch64"], ["--workload", "hmc5000"]])
def test_more_than_one_gpu_exits(extra):
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2", *extra], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""
