"""The generated gfx950 code of the pipelined BF16 GEMM kernel keeps the rules
its inline asm relies on (asm_check.py states them).  CPU only: hipcc
compiles the kernels file to assembly; nothing runs on a GPU.

The checker itself is pinned first: mistakes planted into the real assembly
text (a register touched while its LDS read is in flight, an accumulator read
right after its MFMA, compiler code using M0) must each be reported.
"""
import os
import shutil
import subprocess

import pytest

import asm_check

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mpstools_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("asm") / "kernels.s"
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "--offload-arch=gfx950", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-O3", "-std=c++17", "--cuda-device-only", "-S", "-x", "hip",
           os.path.join(CSRC, "mpsa_kernels.hip"), "-I", CSRC, "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text()


def _pipe_functions(asm):
    fns = list(asm_check.functions(asm))
    # plain and split forms, each with row-major (bf16) and column-major (double) output
    assert len(fns) == 8, [n for n, _ in fns]
    return fns


def test_checker_reports_planted_mistakes(asm):
    name, lines = next((n, ls) for n, ls in _pipe_functions(asm) if "Lb1E" in n)
    assert asm_check.check(lines) == []
    # a VALU write to the destination of an LDS read issued just before
    k = next(i for i, l in enumerate(lines) if l.strip().startswith("ds_read_b128") and "offset" in l)
    dest = lines[k].split()[1].rstrip(",")
    first = int(dest.strip("v[]").split(":")[0])
    bad = lines[:k + 1] + ["\tv_mov_b32_e32 v%d, 0" % first] + lines[k + 1:]
    assert any("still in flight" in e for e in asm_check.check(bad))
    # an accumulator read right after the MFMA that wrote it
    k = next(i for i, l in enumerate(lines) if l.strip().startswith("v_mfma"))
    acc = lines[k].split()[1].rstrip(",")
    first = int(acc.strip("a[]").split(":")[0])
    bad = lines[:k + 1] + ["\tv_accvgpr_read_b32 v0, a%d" % first] + lines[k + 1:]
    assert any("after an MFMA wrote it" in e for e in asm_check.check(bad))
    # compiler-generated code (outside the asm markers) using M0
    k = next(i for i, l in enumerate(lines) if l.strip().startswith(";;#ASMEND"))
    bad = lines[:k + 1] + ["\ts_mov_b32 m0, s0"] + lines[k + 1:]
    assert any("M0" in e for e in asm_check.check(bad))


def test_pipe_kernels_keep_async_read_and_mfma_rules(asm):
    for name, lines in _pipe_functions(asm):
        errs = asm_check.check(lines)
        assert errs == [], (name, errs[:5])


def test_pipe_kernels_do_not_spill(asm):
    """A spill would move registers behind the asm statements' backs; the
    kernels use no scratch memory."""
    import re
    for m in re.finditer(r"\.amdhsa_kernel (\S*gemm_bf16_pipe\S*)\n(.*?)\.end_amdhsa_kernel", asm, re.S):
        scratch = int(re.search(r"\.amdhsa_private_segment_fixed_size\s+(\d+)", m.group(2)).group(1))
        lds = int(re.search(r"\.amdhsa_group_segment_fixed_size\s+(\d+)", m.group(2)).group(1))
        assert scratch == 0, m.group(1)
        assert lds <= 160 * 1024, m.group(1)
