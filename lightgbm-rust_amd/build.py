"""Builds liblgbm_hip.so (HIP kernels for gfx950 + host driver + C ABI) in-tree.

    python lightgbm-rust_amd/build.py
"""
import os
import re
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
# -packed-fp32-ops: the device pass emits no packed FP32 adds/multiplies/FMAs
# (the host pass warns that it does not know the feature; the device pass
# honours it). check_no_packed_fp32 verifies the result on every build.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
LLVM = os.environ.get("LGBMHIP_LLVM", "/opt/rocm/llvm/bin")
PACKED_FP32 = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def check_no_packed_fp32(obj):
    """Disassembles the gfx950 code object embedded in `obj` (a hipcc -c
    output) and fails if any packed FP32 add, multiply or FMA is in it."""
    fat, co = obj + ".fatbin", obj + ".gfx950.co"
    try:
        subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", obj, fat], check=True)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--type=o", f"--input={fat}",
                        f"--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}", f"--output={co}", "--unbundle"], check=True)
        asm = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", co], check=True, capture_output=True,
                             text=True).stdout
    finally:
        for f in (fat, co):
            if os.path.exists(f):
                os.remove(f)
    if not asm.strip():
        raise RuntimeError(f"no device code disassembled from {obj}")
    hits = [ln.strip() for ln in asm.splitlines() if PACKED_FP32.search(ln)]
    if hits:
        raise RuntimeError(f"{obj}: {len(hits)} packed FP32 instructions in device code, e.g. {hits[0]}")
    print(f"{os.path.basename(obj)}: no packed FP32 instructions in {asm.count(chr(10))} lines of gfx950 code")


def build():
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_bin = os.path.join(OUT_DIR, "binning.o")
    obj_eng = os.path.join(OUT_DIR, "engine.o")
    obj_met = os.path.join(OUT_DIR, "metrics.o")
    run(["g++", *COMMON, "-Wall", "-c", os.path.join(CSRC, "binning.cpp"), "-o", obj_bin, "-pthread"])
    run([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "engine.hip"), "-o", obj_eng])
    run([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "metrics.hip"), "-o", obj_met])
    check_no_packed_fp32(obj_eng)
    check_no_packed_fp32(obj_met)
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", *DEVICE, obj_eng, obj_met, obj_bin, "-o", tmp, "-pthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
