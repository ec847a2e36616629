"""Builds tests/dropin/cholesky_dropin.cpp with g++ against the C-ABI library, as INTEGRATION.md's
link line does (used by the drop-in tests; the binary goes to a temporary directory)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def build(out_dir: Path) -> Path:
    libdir = ROOT / "tasktorrent_amd" / "_lib"
    exe = Path(out_dir) / "cholesky_dropin"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", str(ROOT / "include"),
           str(ROOT / "tests" / "dropin" / "cholesky_dropin.cpp"), "-o", str(exe),
           "-L", str(libdir), "-ltasktorrent_amd", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in build failed:\n{r.stderr}")
    return exe
