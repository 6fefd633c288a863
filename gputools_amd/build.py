"""Builds libgputools_amd.so (gfx950 HIP kernels + the C-ABI) in-tree.

    python -m gputools_amd.build          # or gputools_amd.build.build()

Every .hip / .cpp file under csrc/ is compiled by hipcc for gfx950 into
csrc/_build/*.o and linked into gputools_amd/libgputools_amd.so.  The library
depends on the HIP runtime only (no torch).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libgputools_amd.so")
# Peak-rate probes for the measurement figures (bench.py --calibrate-peaks);
# a library of their own, not linked into the distance library.
CALIB_SRC = os.path.join(HERE, "calib", "gt_calib.hip")
CALIB_LIB = os.path.join(HERE, "libgt_calib.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

# -packed-fp32-ops off: the compiler must not form v_pk_{add,mul,fma}_f32 from
# scalar or float2 code.
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(os.path.dirname(HERE), "include")]
# Code object v5: loadable by the HIP runtime torch bundles as well as /opt/rocm's.
DEVICE = ["--offload-arch=" + ARCH, "-mcode-object-version=5", "-Xclang", "-target-feature",
          "-Xclang", "-packed-fp32-ops", "-ffp-contract=fast", "-munsafe-fp-atomics"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + DEVICE
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-mcode-object-version=5", "-shared", "-fPIC",
           "-o", tmp] + objs + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    build_calib()
    if verbose:
        print("built", LIB, "and", CALIB_LIB)
    return LIB


def build_calib() -> str:
    tmp = CALIB_LIB + ".tmp"
    cmd = [HIPCC] + COMMON + DEVICE + ["-shared", CALIB_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (CALIB_SRC, " ".join(cmd), r.stderr))
    os.replace(tmp, CALIB_LIB)
    return CALIB_LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
