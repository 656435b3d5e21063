"""Build the engine in-tree for gfx950.

    python -m branson_amd.build          # libbranson_hip.so + host driver
    python -m branson_amd.build --all    # also the oracle (test infrastructure)

Outputs (git-ignored, shipped with the tree):
    branson_amd/libbranson_hip.so  -- HIP kernels + C-ABI (include/branson_hip.h)
    branson_amd/branson_hip        -- command-line driver for Branson XML decks
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libbranson_hip.so")
EXE = os.path.join(PKG, "branson_hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BH_OFFLOAD_ARCH", "gfx950")

# Device code replays IEEE double arithmetic: no contraction into FMA.
# Packed FP32 instructions are disabled (they have returned wrong results
# intermittently on this hardware).  FP atomics map to the hardware
# global_atomic_add_f64 / ds_add_f64 instructions.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              "-Wno-unused-parameter"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_engine():
    srcs = [os.path.join(CSRC, "bh_engine.hip")]
    host_srcs = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
                       if f.endswith(".cpp")) if os.path.isdir(os.path.join(CSRC, "host")) else []
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs, *host_srcs, "-I", os.path.join(ROOT, "include")])
    os.replace(tmp, LIB)
    main = os.path.join(CSRC, "host_main", "branson_main.cpp")
    if os.path.exists(main):
        tmp = EXE + ".tmp"
        _run(["g++", *HOST_FLAGS, "-o", tmp, main, "-I", os.path.join(ROOT, "include"),
              "-I", os.path.join(CSRC, "host"), f"-L{PKG}", "-lbranson_hip",
              f"-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, EXE)


def build_oracle(ref: bool):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    if ref and os.path.isdir("/root/reference/src/random123"):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--all", action="store_true", help="also build the oracle")
    a = ap.parse_args(argv)
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        sys.exit(f"hipcc not found at {HIPCC}")
    build_engine()
    if a.all:
        build_oracle(ref=True)


if __name__ == "__main__":
    main()
