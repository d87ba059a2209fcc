"""In-tree build of libtjsq_hip.so for gfx950 (hipcc, no JIT cache)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "tjsq_plan.cpp"), os.path.join(HERE, "csrc", "tjsq_update.cpp"),
           os.path.join(HERE, "csrc", "tjsq_packed.cpp"),
           os.path.join(HERE, "csrc", "tjsq_hip.hip")]
OUT = os.path.join(HERE, "libtjsq_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc"),
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    # rocSOLVER (with rocBLAS) diagonalises the density-matrix sectors
    libs = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-lpthread", "-Wl,-rpath,/opt/rocm/lib"]
    cmd = [hipcc] + HIPCC_FLAGS + SOURCES + libs + ["-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed to build libtjsq_hip.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
