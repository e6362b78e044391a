"""Builds liblgbm_hip.so (HIP kernels for gfx950 + host driver + C ABI) in-tree.

    python lightgbm-rust_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "liblgbm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LGBMHIP_ARCH", "gfx950")

# -ffp-contract=off: every a*b+c stays two roundings, as the reference computes
# it (split gains and leaf outputs are compared bit for bit).
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build():
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_bin = os.path.join(OUT_DIR, "binning.o")
    obj_eng = os.path.join(OUT_DIR, "engine.o")
    obj_met = os.path.join(OUT_DIR, "metrics.o")
    run(["g++", *COMMON, "-Wall", "-c", os.path.join(CSRC, "binning.cpp"), "-o", obj_bin, "-pthread"])
    run([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "engine.hip"), "-o", obj_eng])
    run([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "metrics.hip"), "-o", obj_met])
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", *DEVICE, obj_eng, obj_met, obj_bin, "-o", tmp, "-pthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
