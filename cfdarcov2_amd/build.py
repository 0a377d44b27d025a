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
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcfdarco_hip.so")
CPP = os.path.join(HERE, "cpp")
FVM_LIB = os.path.join(HERE, "libcfdarco_fvm3d.so")
BUILD_DIR = os.path.join(HERE, "_build")
REFERENCE = "/root/reference"
# the reference's example drivers, compiled unchanged against the operator API
REFERENCE_DRIVERS = {
    "cfdARCO_heat": "examples/main_3d_heat.cpp",
    "cfdARCO_wave": "examples/main_3d_wave.cu",
    "cfdARCO_gauss_pulse": "examples/main_3d_wave_gauss_pulse.cu",
    "cfdARCO_sod_shock": "examples/main_3d_sod_shock.cpp",
    "cfdARCO_euler": "examples/main_3d.cpp",
}
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


# host C++ of the operator API and its drivers (hipcc, so HIP kernels in
# driver sources compile too)
CPP_FLAGS = ["--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed building {what}")


def build_fvm3d(verbose: bool = False) -> str:
    """libcfdarco_fvm3d.so: the C++ operator API (Variable, Mesh3D, Equation,
    KernelBuilder) on top of libcfdarco_hip.so."""
    srcs = sorted(glob.glob(os.path.join(CPP, "src", "*.cpp")))
    tmp = FVM_LIB + ".tmp"
    cmd = [HIPCC, *CPP_FLAGS, "-x", "hip", *srcs, "-shared", "-o", tmp, f"-I{CPP}/include", f"-L{HERE}",
           "-lcfdarco_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM_LLVM_LIB}", "-fopenmp"]
    _run(cmd, "libcfdarco_fvm3d.so", verbose)
    os.replace(tmp, FVM_LIB)
    return FVM_LIB


def build_driver(src: str, out: str, verbose: bool = False, extra=()) -> str:
    """Compile a driver (reference example or test program) against the operator API."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cmd = [HIPCC, *CPP_FLAGS, "-x", "hip", src, *extra, "-o", out, f"-I{CPP}/include", f"-L{HERE}",
           "-lcfdarco_fvm3d", "-lcfdarco_hip", f"-Wl,-rpath,{HERE}", f"-Wl,-rpath,{ROCM_LLVM_LIB}"]
    _run(cmd, os.path.basename(out), verbose)
    return out


def build_reference_drivers(verbose: bool = False) -> list:
    """The reference's drivers, unchanged, when the reference tree is present."""
    built = []
    if not os.path.isdir(REFERENCE):
        return built
    for name, rel in REFERENCE_DRIVERS.items():
        src = os.path.join(REFERENCE, rel)
        if os.path.exists(src):
            built.append(build_driver(src, os.path.join(BUILD_DIR, "drivers", name), verbose))
    return built


def build_test_programs(verbose: bool = False) -> list:
    """C++ parity programs under tests/cpp (written against the operator API)."""
    built = []
    for src in sorted(glob.glob(os.path.join(ROOT, "tests", "cpp", "*.cpp"))):
        name = os.path.splitext(os.path.basename(src))[0]
        built.append(build_driver(src, os.path.join(BUILD_DIR, "tests", name), verbose))
    return built


def build(verbose: bool = False) -> str:
    """Compile the library; returns the path of the shared object.  Each
    translation unit compiles to its own object in parallel, then one link."""
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(BUILD_DIR, "obj")
    os.makedirs(objdir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        return obj, res

    workers = max(1, min(len(sources()), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(workers) as pool:
        results = list(pool.map(compile_one, sources()))
    failed = [r for _, r in results if r.returncode != 0]
    if failed:
        for r in failed:
            sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libcfdarco_hip.so")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *[o for o, _ in results],
           f"-Wl,-rpath,{ROCM_LLVM_LIB}", "-lhiprtc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libcfdarco_hip.so")
    os.replace(tmp, LIB)
    return LIB


def build_bench_programs(verbose: bool = False) -> list:
    """C++ benchmark helpers under cpp/bench (written against the operator API)."""
    built = []
    for src in sorted(glob.glob(os.path.join(CPP, "bench", "*.cpp"))):
        name = os.path.splitext(os.path.basename(src))[0]
        built.append(build_driver(src, os.path.join(BUILD_DIR, "bench", name), verbose))
    return built


STREAM_PROBE = os.path.join(BUILD_DIR, "bench", "libstream_probe.so")


def build_stream_probe(verbose: bool = False) -> str:
    """The streaming-bandwidth probe the benchmark reports its ceiling from."""
    os.makedirs(os.path.dirname(STREAM_PROBE), exist_ok=True)
    src = os.path.join(CPP, "bench", "stream_probe.hip")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", STREAM_PROBE, src]
    _run(cmd, "the stream probe", verbose)
    return STREAM_PROBE


def build_all(verbose: bool = False):
    out = [build(verbose), build_fvm3d(verbose)]
    out += build_bench_programs(verbose)
    out.append(build_stream_probe(verbose))
    out += build_test_programs(verbose)
    out += build_reference_drivers(verbose)
    return out


if __name__ == "__main__":
    for path in build_all(verbose="-v" in sys.argv):
        print(path)
