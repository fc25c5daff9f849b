"""Field arithmetic as compiled for gfx950 against the same functions compiled
for the host.

The device form of the 128-bit reduction is inline assembly (the borrow of a
64-bit subtract and the carry-out of v_mad_u64_u32 as lane masks); the host
form is the C++ the CPU tests pin against Python integers.  `gnt_calib field`
runs add, sub, mul and the reduction on every pair of 39 edge words (around 0,
2^32 - 1, 2^32, 2^63, p and 2^64, and words whose top half equals or exceeds
the bottom word, the borrow cases) and on 2^20 random pairs (with forced
halves), x * 2^e for every e < 192 as the register DFTs compile it, and radix-8
and radix-16 register DFTs, and counts any difference from the host forms.
"""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_device_field_arithmetic_matches_host_forms():
    exe = os.path.join(ROOT, "pi_amd", "gnt_calib")
    assert os.path.exists(exe), "pi_amd/gnt_calib is not built"
    print("gnt_calib field", flush=True)
    r = subprocess.run([exe, "field"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["calib"] == "field"
    assert res["ops_differing"] == 0 and res["pow2_differing"] == 0 and res["dft_differing"] == 0, res
    assert res["ops_checked"] >= 6 * (1 << 20) and res["pow2_checked"] >= 192 * 65536, res
