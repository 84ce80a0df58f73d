"""Builds libvrs.so (the C-ABI of include/vrs.h) in-tree for gfx950 with hipcc.

    python -m viss_recon_shade_amd.build

Objects go to viss_recon_shade_amd/_build/, the library to
viss_recon_shade_amd/libvrs.so.  Every compile of device code carries
-Xclang -target-feature -Xclang -packed-fp32-ops (no packed FP32 VALU ops).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libvrs.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["capi.hip", "kernels.hip", "mfma.hip", "refcompat.hip", "grid.hip", "host_sky.cpp", "io.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          "-Wall", "-Wno-unused-function"]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]
# Per-source device flags.  kernels.hip keeps its per-visibility accumulators
# in unrolled register arrays updated under wave-uniform branches; promoting
# those arrays to vector values makes every branch copy the whole array.
# Machine scheduler strategies, chosen by timing on an MI355X (c2 plain, c2
# FP64-phase and c3 64-map kernels; every default-strategy kernel was 0.4-2.4 %
# slower): max-ilp for the VALU-bound DFT kernels, iterative-ilp for the
# matrix-core kernels (four waves per SIMD with long dependent MFMA chains:
# instruction-level parallelism inside a wave is what overlaps their VALU work
# with those chains).
EXTRA = {"kernels.hip": ["-mllvm", "-disable-promote-alloca-to-vector",
                         "-mllvm", "--amdgpu-sched-strategy=max-ilp"],
         "mfma.hip": ["-mllvm", "--amdgpu-sched-strategy=iterative-ilp"]}


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = ["g++"] + COMMON + ["-c", path, "-o", out]
    else:
        cmd = [HIPCC] + COMMON + DEVICE + EXTRA.get(src, []) + ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
