"""Write the end-to-end pi fixtures (test infrastructure; run in the build container).

Builds pi_golden.cpp (GMP), computes pi by the Chudnovsky series in decimal
fixed point and by the Gauss-Legendre AGM in binary fixed point, requires the
two to agree wherever both are run, and records:
  pi_digits.json  SHA-256 and byte length of the "3." + (D-1) digits + "\\n"
                  files (crates/pi-core/src/output.rs:33-61) for each D
  pi_10000.txt    the 10,000-digit file itself

Usage: python tests/golden/make_pi_golden.py [--max-digits 100000000]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
GMP = "/usr/lib/x86_64-linux-gnu/libgmp.so.10"


def build(out_dir: str) -> str:
    exe = os.path.join(out_dir, "pi_golden")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "pi_golden.cpp"), "-o", exe, GMP], check=True)
    return exe


def run(exe: str, algo: str, digits: int, path: str) -> float:
    t0 = time.time()
    subprocess.run([exe, algo, str(digits), path], check=True)
    return time.time() - t0


def digest(path: str):
    h = hashlib.sha256()
    n = 0
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
            n += len(chunk)
    return h.hexdigest(), n


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--max-digits", type=int, default=100_000_000)
    ap.add_argument("--gl-max-digits", type=int, default=10_000_000)
    args = ap.parse_args()
    sizes = [d for d in (10_000, 100_000, 1_000_000, 10_000_000, 100_000_000) if d <= args.max_digits]
    entries = {}
    with tempfile.TemporaryDirectory() as tmp:
        exe = build(tmp)
        for d in sizes:
            p1 = os.path.join(tmp, f"c{d}.txt")
            t1 = run(exe, "chudnovsky", d, p1)
            h1, n1 = digest(p1)
            e = {"sha256": h1, "bytes": n1, "chudnovsky_s": round(t1, 2)}
            if d <= args.gl_max_digits:
                p2 = os.path.join(tmp, f"g{d}.txt")
                t2 = run(exe, "gauss-legendre", d, p2)
                h2, _ = digest(p2)
                if h1 != h2:
                    raise SystemExit(f"Chudnovsky and Gauss-Legendre disagree at {d} digits")
                e["gauss_legendre_s"] = round(t2, 2)
                e["cross_checked"] = "gauss-legendre"
                os.remove(p2)
            if d == 10_000:
                with open(p1, "rb") as f, open(os.path.join(HERE, "pi_10000.txt"), "wb") as g:
                    g.write(f.read())
            os.remove(p1)
            entries[str(d)] = e
            print(d, e, flush=True)
    with open(os.path.join(HERE, "pi_digits.json"), "w") as f:
        json.dump({"format": "3. + (D-1) digits + newline", "generator": "tests/golden/pi_golden.cpp",
                   "digits": entries}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
