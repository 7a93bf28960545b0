"""Builds tuvok_amd/libtvkhip.so (HIP kernels + C ABI) for gfx950 with hipcc.

Run as ``python -m tuvok_amd.build`` or through ``__graft_entry__.build()``.
Objects go to tuvok_amd/_build/; the shared library lands next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtvkhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TVK_OFFLOAD_ARCH", "gfx950")

# (source, extra flags).  The march kernel must not contract a*b+c into FMAs:
# its float operations mirror the host planner's and the reference restatement's.
SOURCES = [
    ("kernels/march.hip", ["-ffp-contract=off"]),
    ("kernels/bricking.hip", []),
    ("kernels/generate.hip", []),
    ("kernels/stepnet.hip", []),
    ("kernels/leap.hip", ["-ffp-contract=off"]),
    ("kernels/display.hip", ["-ffp-contract=off"]),
    ("host/host.cpp", ["-ffp-contract=off"]),
    ("api.cpp", ["-ffp-contract=off"]),
    ("batch.cpp", ["-ffp-contract=off"]),
    ("stepnet_api.cpp", []),
    ("leap_api.cpp", ["-ffp-contract=off"]),
    ("noalloc.hip", ["-ffp-contract=off"]),
    ("pool.cpp", ["-ffp-contract=off"]),
    ("host/tf2d.cpp", ["-ffp-contract=off"]),
    ("host/uvf.cpp", []),
]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # never emit packed fp32 VALU ops (v_pk_{add,mul,fma}_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-result",
]


def _compile(item):
    src, extra = item
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    cmd = [HIPCC, "-c", path, "-o", obj] + COMMON + extra
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip", "-c", path, "-o", obj] + COMMON + extra
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
