"""Build the HIP shared library ``jp2lam_amd/libjp2lam_amd.so`` for gfx950.

Usage: ``python -m jp2lam_amd.build`` (or ``build()`` from Python).  Every
source is compiled by hipcc with the flags the kernels rely on:

* ``-ffp-contract=off`` -- no FMA contraction, so the f32/f64 arithmetic
  follows the reference's evaluation order bit for bit;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops`` -- no packed FP32
  instructions in the device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libjp2lam_amd.so"
BUILD_DIR = PKG / "_build"

SOURCES = ["kernels_dwt.hip", "kernels_t1.hip", "kernels_pcrd.hip", "kernels_metrics.hip", "host.cpp", "engine.cpp", "queue.cpp", "server.cpp", "api.cpp"]
ARCH = os.environ.get("JP2LAM_AMD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build jp2lam_amd)")


def _flags(src: str) -> list[str]:
    f = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
    if src.endswith(".cpp"):
        f += ["-x", "hip", "-pthread"]
    return f


def build(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    BUILD_DIR.mkdir(exist_ok=True)

    def compile_one(src: str) -> Path:
        obj = BUILD_DIR / (src + ".o")
        cmd = [hipcc, *_flags(src), "-c", str(CSRC / src), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
        if verbose:
            print("compiled", src, flush=True)
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc, "-shared", "-fPIC", "-pthread", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
