"""The kernels that read other lanes' registers (DPP lane shifts for z
neighbours, readlane and ballots for the sensors, shuffles in reductions)
must not spill registers: a spilled
value is reloaded only in the lanes active at the reload, so a later read of
an inactive lane's copy sees a stale value.  (An FP64 8-row build of the
transposed row step spilled and did exactly that: it reloaded a sensor's
trace gradient inside the branch of the lane holding the sensor's cell, then
read it from the sensor's own lane, and its gradient differed from the other
forms'.)  This test compiles every
translation unit that includes such kernels for gfx950 and checks that each
kernel's private segment is empty.  It needs hipcc, not a GPU."""
import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

from cfdarcov2_amd import build

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cfdarcov2_amd", "csrc")
CROSS_LANE_UNITS = ["adj_row.hip", "adjoint.hip", "batch.hip", "euler.hip", "ewise.hip",
                    "heat_multi.hip", "heat_row.hip", "wave_row.hip", "wave_row2.hip"]


def _have_hipcc():
    return os.path.exists(build.HIPCC) or shutil.which("hipcc") is not None


def _units_including_cross_lane_code():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if not f.endswith(".hip"):
            continue
        src = open(os.path.join(CSRC, f)).read()
        if re.search(r'#include "row2?\.hpp"|update_dpp|readlane|__shfl|__ballot', src):
            out.append(f)
    return out


def test_cross_lane_unit_list_is_complete():
    assert _units_including_cross_lane_code() == CROSS_LANE_UNITS


@pytest.mark.skipif(not _have_hipcc(), reason="hipcc not available")
def test_cross_lane_kernels_do_not_spill():
    with tempfile.TemporaryDirectory() as tmp:
        def compile_unit(f):
            out = os.path.join(tmp, f + ".s")
            r = subprocess.run([build.HIPCC, *build.HIP_FLAGS, "--cuda-device-only", "-S", "-o", out,
                                os.path.join(CSRC, f)], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-2000:]
            return f, open(out).read()

        with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
            asm = list(ex.map(compile_unit, CROSS_LANE_UNITS))
    spilled = []
    kernels = 0
    for f, text in asm:
        for m in re.finditer(r"\.set (\S+)\.private_seg_size, (\d+)", text):
            kernels += 1
            if int(m.group(2)) > 0:
                spilled.append((f, m.group(1), int(m.group(2))))
    assert kernels > 50
    assert not spilled, spilled
