"""Build libctraj.so in-tree with hipcc for gfx950."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
INCLUDE = PKG.parent.parent / "include"
OUT = PKG / "libctraj.so"

SOURCES = ["ctraj.hip", "batch.hip", "adjoint.hip", "sort.hip", "resample.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",  # every fused multiply-add in the kernels is written out explicitly
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(out: Path = OUT, verbose: bool = False) -> Path:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc, *FLAGS, "-I", str(INCLUDE), "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
