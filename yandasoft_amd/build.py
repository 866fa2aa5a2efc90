"""Build recipe for libygrid.so (HIP for gfx950, in-tree).

    python -m yandasoft_amd.build          # or:  python yandasoft_amd/build.py

Every translation unit is compiled by hipcc for gfx950 and linked with hipFFT into
``yandasoft_amd/libygrid.so``.  Packed-FP32 instructions are disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``) and float atomics compile to the
hardware instructions (``-munsafe-fp-atomics``).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libygrid.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

SOURCES = [
    "kernels_grid.hip",
    "kernels_small.hip",
    "kernels_fft.hip",
    "kernels_cf.hip",
    "kernels_awp.hip",
    "kernels_mlp.hip",
    "kernels_uvw.hip",
    "kernels_snapshot.hip",
    "spheroidal.cpp",
    "gridders.cpp",
    "awproject.cpp",
    "batch.cpp",
    "learned_cf.cpp",
    "imaging.cpp",
    "polconv.cpp",
    "snapshot.cpp",
    "noalloc.cpp",
    "ygrid_api.cpp",
]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(HERE, "..", "include"),
]


def _compile(src: str, objdir: str) -> str:
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L" + os.path.join(ROCM, "lib"), "-lhipfft", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
