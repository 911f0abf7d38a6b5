"""Compile the gfx950 HIP kernels into pyseqm_amd/libpyseqm_hip.so (in-tree, no JIT cache).

Usage: python -m pyseqm_amd.csrc.build   (or via __graft_entry__.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
SOURCES = ["pairs.hip", "fock.hip", "scf.hip", "sp2.hip", "xl.hip", "eigh.hip", "layout.hip", "ksa.hip"]
OUT = os.path.join(PKG, "libpyseqm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         # packed-FP32 VALU forms have given wrong results on these parts: keep them off
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-Wno-unused-result"]


def build(verbose=True):
    srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
