"""Build liborb_gpu.so (gfx950 kernels + C ABI) in-tree with hipcc.

Flags:
  -ffp-contract=off   float statements are evaluated unfused, exactly as
                      written, on the host and on the device
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in device code
  -mllvm -amdgpu-mfma-vgpr-form
                      MFMA accumulators in VGPRs, so the Hamming kernel's
                      epilogue reads them without v_accvgpr_read copies
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "liborb_gpu.so")
SOURCES = ["orb_kernels.hip", "orb_bow.hip", "orb_mappoint.hip", "orb_color.hip", "orb_winmatch.hip",
           "orb_frame.hip", "orb_runtime.cpp", "orb_serve.cpp", "orb_vocabulary.cpp", "orb_rectify.cpp", "orb_winmatch.cpp"]
ARCH = os.environ.get("ORB_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def command(out=LIB):
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-amdgpu-mfma-vgpr-form",
            "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", CSRC, "-o", out] + \
           [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building liborb_gpu.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
