"""The host driver's multi-threaded shifts, sums, differences and word
multiplies (pi_amd/csrc/par_mpz.hpp) give GMP's results bit for bit
(tests/emu/par_mpz_check.cpp: every operation on fixed random, all-ones and
sparse numbers of 1 to 100,003 limbs, both signs, operands aliasing the
result)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GMP = "/usr/lib/x86_64-linux-gnu/libgmp.so.10"


@pytest.mark.skipif(not os.path.exists(GMP), reason="libgmp.so.10 not present")
def test_par_mpz_matches_gmp():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "emu"), os.path.join(HERE, "emu", "par_mpz_check")],
                   check=True)
    r = subprocess.run([os.path.join(HERE, "emu", "par_mpz_check")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "fails 0" in r.stdout, r.stdout + r.stderr
