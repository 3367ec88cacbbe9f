"""Build the HIP library in-tree (gfx950) and the host-only small-LA test library.

    python -m titus_amd.build          # or __graft_entry__.build()

Outputs (git-ignored, shipped with the tree):
    titus_amd/libtitus_hip.so   -- the C-ABI of include/titus_hip.h
    titus_amd/_smallla_host.so  -- host dense LA with a C test interface
    titus_amd/hbm_ceiling       -- the measured HBM ceilings (bench.py runs it)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtitus_hip.so")
SMALLLA = os.path.join(HERE, "_smallla_host.so")
HBM_CEILING = os.path.join(HERE, "hbm_ceiling")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core accumulators in ordinary VGPRs: without it the loop-carried
    # accumulators are copied to and from AGPRs around every MFMA loop trip
    # (and the copies wait for the MFMA pipeline to drain)
    # (an AMDGPU backend option; the host pass registers it too and ignores it)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument",
    # host-side small algebra: AVX2 + FMA (every x86-64 server CPU of an
    # MI355X node has them)
    "-Xarch_host", "-mavx2", "-Xarch_host", "-mfma",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return res


def build(verbose=False):
    hipcc = _hipcc()
    srcs = [os.path.join(CSRC, f) for f in ("kernels.hip", "angular.hip", "solver.cpp", "adjoint.cpp",
                                            "fullpn.cpp", "primitives.hip", "smallla.cpp")]
    tmp = LIB + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-shared", "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp])
    os.replace(tmp, LIB)
    tmp = SMALLLA + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared",
          os.path.join(CSRC, "smallla.cpp"), os.path.join(CSRC, "smallla_capi.cpp"), "-o", tmp])
    os.replace(tmp, SMALLLA)
    tmp = HBM_CEILING + ".tmp"
    _run([hipcc, *HIP_FLAGS, os.path.join(CSRC, "hbm_ceiling.hip"), "-o", tmp])
    os.replace(tmp, HBM_CEILING)
    if verbose:
        print("built", LIB, SMALLLA, "and", HBM_CEILING)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
