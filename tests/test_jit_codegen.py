"""KernelBuilder's generated sources compile with hipRTC for gfx950 (host only,
no device): tests/cpp/jit_codegen.cpp emits the fused kernels of the
operator-API test systems, and the kernels that compute stabilised pointwise
operands, and each is compiled through cfd_program_compile_only."""
import glob
import os
import subprocess
import tempfile

import pytest

from cfdarcov2_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "cfdarcov2_amd", "_build", "tests", "jit_codegen")


def test_generated_kernels_compile():
    if not os.path.exists(PROG):
        pytest.fail(f"{PROG} not built (run __graft_entry__.build())")
    with tempfile.TemporaryDirectory() as tmp:
        res = subprocess.run([PROG, tmp], capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        names = [l for l in res.stdout.split() if l.startswith("cfd_fused_")]
        operands = [l for l in res.stdout.split() if l.startswith("cfd_operand_")]
        assert len(names) == 12
        # stab_tot(E): a pointwise operand of several operators gets a kernel of its own
        assert len(operands) == 1
        files = sorted(glob.glob(os.path.join(tmp, "*.hip")))
        assert len(files) == len(set(names)) + len(operands)
        for path in files:
            src = open(path).read()
            name = os.path.splitext(os.path.basename(path))[0]
            assert f'extern "C" __global__' in src and name in src
            assert capi.compile_only(src, name) > 0
        # the stabilisation and the stencil-in-stencil are evaluated before the
        # launch and read as buffers; fused stencils run in registers
        srcs = {os.path.basename(p): open(p).read() for p in files}
        assert any("jlapl(" in s and "jgrad(" in s for s in srcs.values())
        assert all("out_[c]" in s for n, s in srcs.items() if n.startswith("cfd_fused_"))
        op_src = srcs[operands[0] + ".hip"]
        assert "res_[c] = " in op_src and "jstab_axis(g, T" not in op_src.split("cfd_operand_")[-1]
