"""Build the gfx950 HIP library liboaph.so in-tree.

hipcc compiles every oap_amd/csrc/*.hip for --offload-arch=gfx950 (cross-
compiles without a GPU) and links them into oap_amd/liboaph.so, the C-ABI
library that include/oaph.h declares.  Run `python -m oap_amd.build`.
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
BUILD = PKG / "_build"
LIB = PKG / "liboaph.so"
ARCH = os.environ.get("OAPH_OFFLOAD_ARCH", "gfx950")

CXXFLAGS = [
    "-O3",
    "-fPIC",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", str(PKG.parent / "include"),
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def _compile(src: Path, hipcc: str) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc, *CXXFLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources under oap_amd/csrc")
    workers = min(len(sources), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda s: _compile(s, hipcc), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-lpthread", "-lz"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    # reference-style C++ callers of the facade header (include/oap/oap.h)
    for example in sorted((PKG.parent / "tests" / "cpp").glob("*.cpp")):
        cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-I",
               str(PKG.parent / "include"), str(example), "-L", str(PKG), "-loaph",
               "-Wl,-rpath,$ORIGIN/../../oap_amd", "-o", str(example.with_suffix(""))]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"{example.name} failed to build:\n{res.stderr}")
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", flush=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
