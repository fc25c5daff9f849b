"""End-to-end Chudnovsky on the GPU: the binary-splitting merges run through the
serving loop (products, sums and differences on device-resident P, Q, T) and
the digits must equal the pinned fixtures byte for byte.

Fixtures (tests/golden/pi_digits.json, pi_10000.txt) come from an independent
GMP program that agrees with a Gauss-Legendre run; the first 100 digits are
the reference's FIRST_100 (crates/pi-core/src/algorithm/util.rs:126-127).
"""
import hashlib
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PI = os.path.join(ROOT, "pi_amd", "pi_compute")
GOLD = os.path.join(ROOT, "tests", "golden")
PLAN_CACHE = os.path.join(ROOT, "pi_amd", "plans_gfx950.json")


def _digest(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def _run(digits, tmp_path, *extra):
    out = tmp_path / f"pi_{digits}.txt"
    print(f"pi_compute {digits} {' '.join(extra)}", flush=True)
    # the tuned plan cache, as the measured runs load it
    r = subprocess.run([PI, str(digits), str(out), "--plans", PLAN_CACHE, *extra], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    stats = json.loads(r.stdout.strip().splitlines()[-1])
    return out, stats


@pytest.mark.parametrize("gpu_min_limbs", ["1", "8", "64"])
def test_pi_10000_gpu_merges(tmp_path, gpu_min_limbs):
    """Almost the whole tree merged on the device (tiny limb threshold)."""
    out, st = _run(10000, tmp_path, "--gpu-min-limbs", gpu_min_limbs)
    assert st["merge"] == "gpu" and st["merge_nodes"] > 0
    with open(os.path.join(GOLD, "pi_10000.txt"), "rb") as f:
        assert out.read_bytes() == f.read()


@pytest.mark.parametrize("digits,gpu_min_limbs,order", [
    (100000, "32", "dataflow"), (1000000, "256", "dataflow"), (10000000, "1024", "dataflow"),
    (100000, "32", "levels"), (1000000, "256", "levels")])
def test_pi_sha256(tmp_path, digits, gpu_min_limbs, order):
    """Merges as children finish (dataflow, waves of mixed depths overlapping
    the host nodes) or a whole tree level at a time: the same digits."""
    gold = json.load(open(os.path.join(GOLD, "pi_digits.json")))["digits"]
    out, st = _run(digits, tmp_path, "--gpu-min-limbs", gpu_min_limbs, "--merge", order)
    assert st["merge"] == "gpu" and st["merge_nodes"] > 0 and st["merge_order"] == order
    assert sum(L["nodes"] for L in st["levels"]) == st["merge_nodes"]
    assert os.path.getsize(out) == gold[str(digits)]["bytes"]
    assert _digest(out) == gold[str(digits)]["sha256"]


@pytest.mark.parametrize("digits,fa_min", [(100000, "16"), (1000000, "256"), (10000000, "1024")])
def test_gauss_legendre_gpu(tmp_path, digits, fa_min):
    """The AGM cross-check with its full-precision products and squares on the GPU."""
    gold = json.load(open(os.path.join(GOLD, "pi_digits.json")))["digits"]
    out, st = _run(digits, tmp_path, "--algorithm", "gauss-legendre", "--fa-min-limbs", fa_min)
    assert st["products"] == "gpu" and st["gpu_products"] > 0
    assert _digest(out) == gold[str(digits)]["sha256"]


@pytest.mark.parametrize("algorithm", ["chudnovsky", "gauss-legendre"])
def test_pi_1e8(tmp_path, algorithm):
    """10^8 digits (7,051,375 series terms for Chudnovsky; 27 AGM iterations at
    full precision for Gauss-Legendre) with the default thresholds, byte for
    byte against the pinned SHA-256."""
    gold = json.load(open(os.path.join(GOLD, "pi_digits.json")))["digits"]["100000000"]
    out, st = _run(100000000, tmp_path, "--algorithm", algorithm)
    assert os.path.getsize(out) == gold["bytes"]
    assert _digest(out) == gold["sha256"]
    print(json.dumps({k: st[k] for k in st if k in ("wall_s", "total_s", "merge", "products")}), flush=True)
