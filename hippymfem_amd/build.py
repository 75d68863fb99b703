"""Builds the native library ``hippymfem_amd/_lib/libhippymfem_amd.so`` in-tree.

One ``hipcc`` link of the HIP kernel files (device code for gfx950 only) and the host C++
files.  Packed FP32 instructions are switched off (``-packed-fp32-ops``) for every HIP file.
Run as ``python -m hippymfem_amd.build``.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libhippymfem_amd.so")
ARCH = os.environ.get("HM_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["hm_element.hip", "hm_mfma.hip", "hm_sumfact.hip", "hm_fused.hip", "hm_scatter.hip",
               "hm_pattern_device.hip", "hm_field.hip"]
CXX_SOURCES = ["hm_abi.cpp", "hm_pattern.cpp"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(OUT_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    cc = hipcc()
    objs = []
    common = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
              "-I", os.path.join(HERE, "..", "include")]
    for name in HIP_SOURCES:
        src = os.path.join(SRC, name)
        if not os.path.exists(src):
            continue
        obj = os.path.join(obj_dir, name + ".o")
        cmd = [cc, "-c", "-x", "hip", "--offload-arch=" + ARCH, *common,
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               src, "-o", obj]
        _run(cmd, verbose)
        objs.append(obj)
    for name in CXX_SOURCES:
        src = os.path.join(SRC, name)
        obj = os.path.join(obj_dir, name + ".o")
        _run([cc, "-c", "-x", "c++", *common, src, "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-o", tmp, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build step failed: %s" % " ".join(cmd))
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
