"""Build the HIP library, the command-line driver and (as test infrastructure)
the CPU oracle, all in-tree.

    python -m mnhack8_amd.build          # library + driver (+ oracle when present)

Outputs: mnhack8_amd/_build/libheat_hip.so (C ABI of include/heat.h),
mnhack8_amd/_build/heat_solver_het (command-line driver).  The GPU code is
compiled for gfx950 only.
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mnhack8_amd")
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(OUT, "libheat_hip.so")
DRIVER = os.path.join(OUT, "heat_solver_het")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # every fused multiply-add is written out as fma(); no implicit contraction
    "-ffp-contract=off",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=ROOT):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    noise = [ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noise:
        print("\n".join(noise[:40]))


def build_library():
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(CSRC, "heat_lib.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), src, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_driver():
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(CSRC, "heat_solver_het.cpp")
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src,
          "-L", OUT, "-lheat_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_oracle():
    """The CPU oracle is test infrastructure; it is built beside the product
    so that tests and the CPU-baseline leg of bench.py find it."""
    mk = os.path.join(ROOT, "oracle", "Makefile")
    if os.path.exists(mk):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])


def build_all():
    build_library()
    build_driver()
    build_oracle()


if __name__ == "__main__":
    build_all()
