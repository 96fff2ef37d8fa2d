"""Build the gfx950 shared library ``frank_amd/libfrank_amd.so`` in tree.

Every source is compiled by hipcc for gfx950 only.  Packed FP32 instructions
are disabled (``-packed-fp32-ops``) for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
BIN = HERE / "bin"
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libfrank_amd.so"
ARCH = "gfx950"

SOURCES = ["kernels.hip", "noalloc.hip", "solve_df.hip", "engine.cpp", "qr.cpp", "small.cpp", "capi.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{HERE.parent / 'include'}", f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> Path:
    out = BUILD / (Path(src).stem + ".o")
    cmd = [hipcc(), "-x", "hip", "-c", str(CSRC / src), "-o", str(out)] + COMMON
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return out


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-o", str(tmp)] + [str(o) for o in objs] + [f"--offload-arch={ARCH}", "-fPIC"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    # the C++ operator-surface driver test links against the library
    src = ROOT / "tests" / "cxx" / "surface_test.cpp"
    if src.exists():
        BIN.mkdir(exist_ok=True)
        exe = BIN / "surface_test"
        cmd = [hipcc(), "-std=c++17", "-O2", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", f"-I{ROOT / 'include'}", str(src), "-o",
               str(exe), f"-L{HERE}", "-lfrank_amd", "-Wl,-rpath,$ORIGIN/.."]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"C++ surface test build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        if verbose:
            print(f"built {exe}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
