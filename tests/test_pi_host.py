"""The Chudnovsky driver's host logic (no GPU): the same split tree merged on
host threads, final assembly and output format, against the pinned fixtures;
and the fixtures themselves against the reference's FIRST_100
(crates/pi-core/src/algorithm/util.rs:126-127)."""
import hashlib
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PI = os.path.join(ROOT, "pi_amd", "pi_compute")
GOLD = os.path.join(ROOT, "tests", "golden")
FIRST_100 = ("3.141592653589793238462643383279502884197169399375105820974944592307816406286208998628034825342117067")


def test_fixture_first_100():
    data = open(os.path.join(GOLD, "pi_10000.txt"), "rb").read()
    assert data.startswith(FIRST_100.encode())
    assert len(data) == 10000 + 2 and data.endswith(b"\n")   # "3." + 9999 digits + "\n"
    gold = json.load(open(os.path.join(GOLD, "pi_digits.json")))["digits"]
    assert hashlib.sha256(data).hexdigest() == gold["10000"]["sha256"]
    # Feynman point: six 9s starting at decimal place 762
    assert data[2 + 761:2 + 767] == b"999999"


def _run(tmp_path, digits, *extra):
    out = tmp_path / "pi.txt"
    r = subprocess.run([PI, str(digits), str(out), *extra], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return out.read_bytes(), json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("digits", [0, 1, 2, 15, 100, 101, 1000])
def test_small_digit_counts(tmp_path, digits):
    data, _ = _run(tmp_path, digits, "--fa-min-limbs", "1e18")
    ref = open(os.path.join(GOLD, "pi_10000.txt"), "rb").read()
    if digits == 0:
        assert data == b"\n"
    elif digits == 1:
        assert data == b"3\n"
    else:
        assert data == ref[:digits + 1] + b"\n"


# Final assembly: "newton" is the driver's own (Newton reciprocal and inverse
# square root, scaled-remainder-tree decimal conversion) with host products;
# "gmp" is GMP's square root, division and radix conversion.
FA = {"newton": ("--fa-min-limbs", "1e18"), "gmp": ("--fa-gmp",)}


@pytest.mark.parametrize("fa", sorted(FA))
@pytest.mark.parametrize("digits,limbs", [(10000, "1"), (10000, "8"), (100000, "64"), (1000000, "512")])
def test_host_merged_tree(tmp_path, digits, limbs, fa):
    data, st = _run(tmp_path, digits, "--host-merge", "--gpu-min-limbs", limbs, *FA[fa])
    assert st["merge"] == "host" and st["merge_nodes"] > 0
    assert st["fa"] == ("gmp" if fa == "gmp" else "host")
    gold = json.load(open(os.path.join(GOLD, "pi_digits.json")))["digits"]
    assert hashlib.sha256(data).hexdigest() == gold[str(digits)]["sha256"]


@pytest.mark.parametrize("digits,threads", [(100000, "4"), (1000000, "7")])
def test_multithreaded_bignum_steps(tmp_path, digits, threads):
    """The final assembly's and the decimal conversion's shifts, sums,
    differences and word multiplies on several threads from 64 limbs up
    (pi_amd/csrc/par_mpz.hpp; the default switches at 131072 limbs), with
    host products: the same digits."""
    gold = json.load(open(os.path.join(GOLD, "pi_digits.json")))["digits"]
    data, st = _run(tmp_path, digits, "--host-merge", "--gpu-min-limbs", "64", "--fa-min-limbs", "1e18",
                    "--par-min-limbs", "64", "--threads", threads)
    assert st["fa"] == "host"
    assert hashlib.sha256(data).hexdigest() == gold[str(digits)]["sha256"]
    data, st = _run(tmp_path, digits, "--algorithm", "gauss-legendre", "--fa-min-limbs", "1e18",
                    "--par-min-limbs", "64", "--threads", threads)
    assert hashlib.sha256(data).hexdigest() == gold[str(digits)]["sha256"]


@pytest.mark.parametrize("digits", [2, 1000, 10000, 100000])
def test_gauss_legendre_host(tmp_path, digits):
    """The AGM cross-check (gauss_legendre.rs:61-166) with host products."""
    data, st = _run(tmp_path, digits, "--algorithm", "gauss-legendre", "--fa-min-limbs", "1e18")
    assert st["algorithm"] == "gauss-legendre" and st["products"] == "host"
    ref = open(os.path.join(GOLD, "pi_10000.txt"), "rb").read()
    if digits <= 10000:
        assert data == (ref[:digits + 1] + b"\n" if digits < 10000 else ref)
    else:
        gold = json.load(open(os.path.join(GOLD, "pi_digits.json")))["digits"]
        assert hashlib.sha256(data).hexdigest() == gold[str(digits)]["sha256"]
