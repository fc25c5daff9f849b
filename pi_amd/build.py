"""Build libgnt.so (the HIP multiply) in-tree with hipcc for gfx950.

Usage: python -m pi_amd.build   (or pi_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgnt.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["gnt_kernels.hip", "gnt_engine.hip", "gnt_batch.hip", "gnt_api.hip", "gnt_dispatch.hip", "gnt_serve.hip",
           "gnt_tune.hip", "gnt_stage.hip", "gnt_noalloc.hip", "gnt_devloop.hip"]

# Packed-FP32 forms are switched off for all device code; the library has no
# floating point, the flag keeps the compiler from introducing any.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


RESOURCES = os.path.join(HERE, "kernel_resources.json")


def parse_resource_remarks(text: str) -> dict:
    """Kernel -> resource figures from clang's kernel-resource-usage remarks."""
    import re
    out, cur = {}, None
    for line in text.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            cur = out.setdefault(m.group(1), {})
            continue
        m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[[^\]]*\])?: (\w+) \[-Rpass-analysis", line)
        if m and cur is not None:
            v = m.group(2)
            cur[m.group(1).strip()] = int(v) if v.isdigit() else v
    return out


def build(verbose: bool = False) -> str:
    objs = []
    resources = {}
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(HERE, "_obj", src.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src))
        objs.append(obj)
    failed = []
    for p, src in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(text)
        resources.update(parse_resource_remarks(text))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    import json
    with open(RESOURCES, "w") as f:
        json.dump(resources, f, indent=0, sort_keys=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # calibration microbenchmarks (HBM copy, modular-arithmetic throughput)
    calib = os.path.join(HERE, "gnt_calib")
    subprocess.run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "gnt_calib.hip"), "-o", calib + ".tmp"], check=True)
    os.replace(calib + ".tmp", calib)
    build_pi_driver()
    return LIB


GMP = "/usr/lib/x86_64-linux-gnu/libgmp.so.10"
PI_BIN = os.path.join(HERE, "pi_compute")


def build_pi_driver() -> str:
    """Host driver of the pi runs (Chudnovsky binary splitting and Gauss-Legendre over libgnt)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-pthread", os.path.join(CSRC, "pi_compute.cpp"),
           "-o", PI_BIN + ".tmp", "-L" + HERE, "-lgnt", GMP, "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    os.replace(PI_BIN + ".tmp", PI_BIN)
    return PI_BIN


if __name__ == "__main__":
    print(build(verbose=True))
