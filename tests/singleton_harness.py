"""Builds tests/c/beast_singleton_harness.c against libmds_hip.so (gcc; the
tests that use it call `harness()`)."""
import subprocess
from pathlib import Path

from massivemds_amd import engine as eng

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "c" / "beast_singleton_harness.c"


def harness(out_dir: Path) -> Path:
    lib = eng.library_path()
    if not lib.exists():
        from massivemds_amd import build

        build.build()
    exe = out_dir / "beast_singleton_harness"
    subprocess.run(["gcc", "-O2", "-std=c99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(SRC),
                    f"-L{lib.parent}", "-lmds_hip", f"-Wl,-rpath,{lib.parent}", "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    return exe
