"""Build recipe for the HIP library (libcfdarco_hip.so, gfx950).

One hipcc invocation over every ``csrc/*.hip`` file, in-tree, so the shared
object travels with the source tree.  Flags:

* ``--offload-arch=gfx950``: MI355X (CDNA4) code objects only;
* ``-ffp-contract=off``: no fused multiply-add contraction, so the kernels keep
  the reference operation order (FP64 results match a CPU evaluation bit for bit);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed-FP32 VALU ops;
* ``-fopenmp`` for the host-side mesh geometry builder.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcfdarco_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LLVM_LIB = "/opt/rocm/lib/llvm/lib"

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fopenmp",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build(verbose: bool = False) -> str:
    """Compile the library; returns the path of the shared object."""
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *sources(),
           f"-Wl,-rpath,{ROCM_LLVM_LIB}", "-lhiprtc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libcfdarco_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
