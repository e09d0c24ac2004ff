"""Builds libmds_hip.so (gfx950 kernels + C ABI) in-tree with hipcc.

Usage: python -m massivemds_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmds_hip.so"
MATHCHECK_LIB = PKG / "libmds_mathcheck.so"
BIN = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    # Packed FP32 VALU forms are kept out of every HIP build.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]

# The library: every .hip file under csrc/ except stand-alone programs.
LIB_SOURCES = ["mds_kernels.hip", "mds_row.hip", "mds_engine.hip", "mds_batch.hip", "mds_hmc.hip", "mds_mlp.hip",
               "mds_singleton.hip", "mds_noalloc.hip"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")


def _compile(src: Path) -> Path:
    out = OBJ / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(out)])
    return out


def build(jobs: int = 8) -> Path:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in LIB_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    # host build of the kernels' arithmetic, for CPU accuracy tests
    mobj = _compile(CSRC / "mds_mathcheck.hip")
    tmp = MATHCHECK_LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", str(mobj), "-o", str(tmp)])
    os.replace(tmp, MATHCHECK_LIB)
    build_cli()
    return LIB


def build_cli() -> Path:
    """mds_benchmark: the command-line benchmark over mds::factory (host C++)."""
    BIN.mkdir(exist_ok=True)
    exe = BIN / "mds_benchmark"
    _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}",
          f"-I{CSRC}", str(CSRC / "mds_factory.cpp"), str(CSRC / "mds_benchmark.cpp"),
          f"-L{PKG}", "-lmds_hip", "-Wl,-rpath,$ORIGIN/..", "-o", str(exe)])
    return exe


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    print(build(args.jobs))


if __name__ == "__main__":
    main()
