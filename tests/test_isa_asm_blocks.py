"""The inline-assembly blocks of the field arithmetic, as the compiler places
them in the transform kernels (CPU test: hipcc cross-compiles for gfx950).

The device reduction's asm blocks write lane masks to scalar registers and one
of them ORs two masks with a scalar instruction, which writes SCC.  The
compiler keeps SCC live across code it believes does not touch it (its own
s_cselect in single-lane regions), so a block that wrote SCC without declaring
it would hand that select a wrong condition.  This test compiles the kernels to
assembly and checks, block by block in straight-line code, that no
instruction reading SCC follows an SCC write made inside an asm block.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pi_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# scalar instructions that do not write SCC (everything else of the s_ form is
# taken to write it, which can only over-report)
NO_SCC_WRITE = ("s_mov", "s_cselect", "s_load", "s_buffer_load", "s_waitcnt", "s_nop", "s_barrier", "s_branch",
                "s_cbranch", "s_setpc", "s_getpc", "s_endpgm", "s_sleep", "s_setprio", "s_movk", "s_cmov",
                "s_mul_i32", "s_mul_hi", "s_set_gpr_idx", "s_memtime", "s_memrealtime", "s_sendmsg", "s_dcache",
                "s_trap", "s_icache", "s_setreg", "s_getreg", "s_delay")
SCC_READERS = ("s_cselect", "s_cbranch_scc", "s_addc", "s_subb", "s_cmov")


def scc_reads_after_asm_writes(asm_text):
    bad, last, inasm = [], None, False
    for line in asm_text.splitlines():
        t = line.strip()
        if re.match(r"^_Z\S+:", line) or t.startswith((".LBB", ".Lfunc_end")):
            last = None
            continue
        if t.startswith(";;#ASMSTART"):
            inasm = True
            continue
        if t.startswith(";;#ASMEND"):
            inasm = False
            continue
        if not t.startswith("s_"):
            continue
        op = t.split()[0]
        if op.startswith(SCC_READERS) and last == "asm":
            bad.append(t)
        if not op.startswith(NO_SCC_WRITE):
            last = "asm" if inasm else "compiler"
    return bad


def test_scc_checker_flags_an_undeclared_write():
    text = "\n".join(["_Zk:", "s_and_b64 s[0:1], s[0:1], exec", ";;#ASMSTART", "s_or_b64 s[2:3], s[2:3], s[4:5]",
                      ";;#ASMEND", "s_cselect_b32 s0, -1, 0"])
    assert scc_reads_after_asm_writes(text) == ["s_cselect_b32 s0, -1, 0"]
    ok = text.replace("s_cselect_b32 s0, -1, 0", "s_cmp_eq_u32 s0, 0\ns_cselect_b32 s0, -1, 0")
    assert scc_reads_after_asm_writes(ok) == []


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("src", ["gnt_kernels.hip", "gnt_engine.hip", "gnt_devloop.hip"])
def test_no_scc_read_after_an_asm_block_writes_it(src, tmp_path):
    out = tmp_path / "k.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--cuda-device-only", "-S", "-o", str(out), os.path.join(CSRC, src)]
    subprocess.run(cmd, check=True, capture_output=True, cwd=CSRC, timeout=600)
    text = out.read_text()
    assert ";;#ASMSTART" in text
    bad = scc_reads_after_asm_writes(text)
    assert not bad, bad[:5]
