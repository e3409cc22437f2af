"""Build libtrixilw_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

    python trixilw.jl_amd/build.py

The library lands next to the Python host package
(trixilw.jl_amd/trixilw_hip/libtrixilw_hip.so) so that it travels with the
source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent
CSRC = PKG_ROOT / "csrc"
OUT = PKG_ROOT / "trixilw_hip" / "libtrixilw_hip.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    # torch's bundled HIP runtime (ROCm 7.0) is the one loaded in-process
    "-mcode-object-version=5",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, out: Path = OUT, defines=()) -> Path:
    srcs = [str(CSRC / "tlw_capi.hip")]
    tmp = Path(out).with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *[f"-D{d}" for d in defines], "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtrixilw_hip.so")
    if verbose and r.stderr.strip():
        # the host pass does not know the packed-fp32 feature; keep other warnings visible
        lines = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    os.replace(tmp, out)
    return Path(out)


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(OUT))
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra preprocessor definition")
    a = ap.parse_args()
    print(build(verbose=True, out=Path(a.out), defines=a.defines))
