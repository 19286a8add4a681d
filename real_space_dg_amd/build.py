"""In-tree build of libwdg.so for gfx950 (hipcc, no cmake)."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libwdg.so"
SOURCES = [CSRC / "wdg.hip", CSRC / "host_setup.cpp"]
HEADERS = [CSRC / "kernels.hpp", CSRC / "theta_fft.hpp", CSRC / "full2d.hpp", CSRC / "full2d_host.hpp",
           CSRC / "host_setup.hpp", CSRC / "reservoir.hpp", PKG.parent / "include" / "wigner_dg.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fopenmp",
    "-Wno-unused-result",
]


def build(verbose: bool = True) -> Path:
    """Compile libwdg.so in place (always recompiles)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-I", str(PKG.parent / "include"), *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
