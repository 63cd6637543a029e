"""Build `pbh_amd/_pm3d_hip.so` for gfx950 with hipcc (in-tree, like pbhgr's `make` of _kernels.so,
reference pbhgr/Makefile:1-8).

    python -m pbh_amd.build            # compile
    python -m pbh_amd.build --check    # print the command only
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc" / "pm3d.hip"
OUT = PKG / "_pm3d_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the FP64 parity kernels must not fuse a*b+c the way numpy never does.
# -packed-fp32-ops off: packed FP32 VALU forms are avoided on this target.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def command(out: Path = OUT):
    return [HIPCC, *FLAGS, f"-I{ROOT / 'include'}", str(SRC), "-o", str(out)]


def build(verbose: bool = True) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    if "--check" in sys.argv:
        print(" ".join(command()))
    else:
        build()
