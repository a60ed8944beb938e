"""Build libpeps_hip.so (hipcc, gfx950) in-tree.

Usage: python -m lgt_amd.build [--jobs N]
Objects go to lgt_amd/_build/, the library to lgt_amd/libpeps_hip.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpeps_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("PEPS_ARCH", "gfx950")

SOURCES = ["kernels.hip", "eigh.hip", "ed.hip", "ctx.cpp", "contract.cpp", "linalg.cpp", "model.cpp",
           "engine.cpp", "mlp.cpp", "dmrg.cpp", "api.cpp"]

CFLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{CSRC}", f"-I{os.path.join(os.path.dirname(HERE), 'include')}",
]


def _hipcc():
    p = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    return p


def _compile(src):
    s = os.path.join(CSRC, src)
    o = os.path.join(OUT, src.rsplit(".", 1)[0] + ".o")
    cmd = [_hipcc(), "-x", "hip", "-c", s, "-o", o] + CFLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return o


def build(jobs: int = 4, verbose: bool = True) -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        f"-L{os.path.join(ROCM, 'lib')}", "-lrocsolver", "-lrocblas",
        f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
