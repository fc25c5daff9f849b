"""The C ABI library builds, loads and exports every entry point include/gnt.h
declares.  No compute calls: these run without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "gnt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gnt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("gnt_init", "gnt_mul", "gnt_sqr", "gnt_mul_mag", "gnt_mul_batch", "gnt_buf_upload",
                 "gnt_mul_buf", "gnt_shutdown", "gnt_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from pi_amd import gnt

    lib = gnt.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", gnt.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (gnt_[a-z0-9_]+)", out))
    assert set(declared()) <= exported
    assert set(gnt.EXPORTS) == set(declared())


def test_library_is_gfx950_code():
    from pi_amd import gnt

    data = open(gnt.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_plan_info_and_errors_without_gpu():
    from pi_amd import gnt

    log_n, radices = gnt.plan_info(1 << 24, 1 << 24)
    assert log_n == 27 and sum(radices) == 27
    # the cap is the reference's: 4(na+nb) <= 2^32 (integer.rs:476-478)
    assert gnt.plan_info(1 << 29, 1 << 29) == (32, [8, 8, 8, 8])
    assert gnt.plan_info((1 << 28) - 7, 1 << 28) == (31, [7, 8, 8, 8])
    with pytest.raises(gnt.GntError) as e:
        gnt.plan_info((1 << 29) + 1, 1 << 29)
    assert e.value.code == gnt.GNT_E_TOO_LARGE
    lib = gnt.load()
    assert lib.gnt_version().startswith(b"gnt")
    # no device in this container: init fails with an error code, never aborts
    h = ctypes.c_void_p()
    rc = lib.gnt_init(ctypes.byref(h), None)
    assert rc in (gnt.GNT_OK, gnt.GNT_E_DEVICE)
    if rc == gnt.GNT_OK:
        lib.gnt_shutdown(h)


def test_plan_overrides_and_cache(tmp_path):
    """gnt_set_plan / gnt_load_plans change the split gnt_plan_info reports
    (host-side state, no device needed); uncovered splits are refused."""
    from pi_amd import gnt
    try:
        assert gnt.plan(24) == (3, [8, 8, 8])          # built-in rule
        gnt.set_plan(24, [7, 7, 10])
        assert gnt.plan(24) == (3, [7, 7, 10])
        with pytest.raises(gnt.GntError):
            gnt.set_plan(24, [6, 8, 10])               # radix 2^6 has no strided tile shape
        with pytest.raises(gnt.GntError):
            gnt.set_plan(24, [8, 8])                   # does not sum to 24
        gnt.set_plan(24)
        assert gnt.plan(24) == (3, [8, 8, 8])
        p = tmp_path / "plans.json"
        p.write_text('{"device": "gfx950",\n "plans": {\n  "17": [9, 8],\n  "22": [8, 7, 7]\n },\n'
                     ' "us": {\n  "17": {"9,8": 1.0}\n }\n}\n')
        gnt.load_plans(str(p))
        assert gnt.plan(17) == (2, [9, 8]) and gnt.plan(22) == (3, [8, 7, 7])
        bad = tmp_path / "bad.json"
        bad.write_text('{"plans": {"17": [5, 12]}}')
        with pytest.raises(gnt.GntError):
            gnt.load_plans(str(bad))
    finally:
        for n in (17, 22, 24):
            gnt.set_plan(n)


def test_transform_kernels_do_not_spill():
    """Every transform kernel keeps its tile within the register budget that
    two 64 KiB tiles per CU allow (8 waves per SIMD, <= 64 VGPRs) and spills
    nothing to scratch.  Figures: the compiler's resource remarks, recorded by
    the build (pi_amd/kernel_resources.json)."""
    import json
    path = os.path.join(ROOT, "pi_amd", "kernel_resources.json")
    if not os.path.exists(path):
        pytest.skip("library built without resource remarks")
    res = json.load(open(path))
    tk = {k: v for k, v in res.items()
          if any(s in k for s in ("k_fwd_pass", "k_inv_pass", "k_last_fused", "k_tpass"))}
    assert len(tk) > 50
    bad = {k: v for k, v in tk.items() if v.get("VGPRs Spill", 0) or v.get("ScratchSize", 0)}
    assert not bad, f"spilling kernels: {sorted(bad)[:5]}"
    # every kernel that runs a full tile holds 8 waves per SIMD
    low = {k: v["Occupancy"] for k, v in tk.items() if v.get("VGPRs", 0) and v.get("Occupancy", 8) < 8}
    assert not low, low
