"""Reductions of the operator API's Eigen subset (cpp/include/cfd_eigen_subset.hpp),
compiled on the host with g++: sum() against an FP64 sum of the same FP32
values, maxCoeff() / minCoeff() against numpy exactly, including lengths
around the eight-lane and 1024-element block edges, values that would stall a
left-to-right FP32 sum, and NaN handling as std::max_element has it."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "cfdarcov2_amd", "cpp", "include")

PROGRAM = r"""
#include <cstdio>
#include <vector>
#include "cfd_eigen_subset.hpp"
int main(int argc, char** argv) {
    FILE* f = std::fopen(argv[1], "rb");
    long long n = 0;
    if (std::fread(&n, sizeof(n), 1, f) != 1) return 2;
    Eigen::Matrix<float, -1, 1> v(n);
    if (n && std::fread(v.data(), sizeof(float), size_t(n), f) != size_t(n)) return 3;
    std::fclose(f);
    Eigen::Matrix<float, -1, 1> w = v;   // copies through the block cache
    std::printf("%.9g %.9g %.9g %.9g\n", double(w.sum()), double(w.maxCoeff()), double(w.minCoeff()),
                double(w.array().sum()));
    return 0;
}
"""


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    d = tmp_path_factory.mktemp("eigen")
    src, exe = d / "red.cpp", d / "red"
    src.write_text(PROGRAM)
    res = subprocess.run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", str(src), "-o", str(exe)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return str(exe), d


def _run(prog, a):
    exe, d = prog
    path = d / "in.bin"
    with open(path, "wb") as f:
        np.array([a.size], dtype=np.int64).tofile(f)
        a.astype(np.float32).tofile(f)
    out = subprocess.run([exe, str(path)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    # nine significant digits round-trip an FP32 value
    return [np.float32(x) for x in out.stdout.split()]


@pytest.mark.parametrize("n", [1, 7, 8, 9, 1023, 1024, 1025, 4099, 1 << 20])
def test_reductions(prog, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n).astype(np.float32) * 10
    s, mx, mn, s2 = _run(prog, a)
    exact = float(np.sum(a, dtype=np.float64))
    assert s == s2
    assert abs(float(s) - exact) <= 1e-6 * float(np.sum(np.abs(a), dtype=np.float64)) + 1e-30
    assert mx == a.max() and mn == a.min()


def test_sum_does_not_stall(prog):
    """2^25 ones: a left-to-right FP32 sum stops at 2^24."""
    a = np.ones(1 << 25, dtype=np.float32)
    s, mx, mn, _ = _run(prog, a)
    assert s == float(1 << 25) and mx == 1.0 and mn == 1.0


def test_nan_handling(prog):
    a = np.arange(20, dtype=np.float32)
    a[5] = np.nan
    _, mx, mn, _ = _run(prog, a)
    assert mx == 19.0 and mn == 0.0          # a NaN after the first element is skipped
    a[0] = np.nan
    _, mx, mn, _ = _run(prog, a)
    assert np.isnan(mx) and np.isnan(mn)      # a leading NaN is kept
