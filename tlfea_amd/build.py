"""Builds libtlfea_hip.so (gfx950) in-tree with explicit hipcc commands.

Every HIP translation unit is compiled to an object with
``--offload-arch=gfx950 -O3 -fPIC -munsafe-fp-atomics`` and packed-FP32 instructions
disabled, then linked into ``tlfea_amd/_build/libtlfea_hip.so``.  The C++ drop-in sources
(``csrc/dropin/*.cpp``) are compiled by hipcc into the same library.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libtlfea_hip.so")
ARCH = os.environ.get("TLFEA_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument",
]


def sources():
    hip = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    cpp = sorted(glob.glob(os.path.join(CSRC, "dropin", "*.cpp")))
    return hip, cpp


def _compile(src: str) -> str:
    obj = os.path.join(OUT, "obj", os.path.basename(src) + ".o")
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    lang = ["-x", "hip"]
    cmd = [HIPCC, *lang, *COMMON_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


DRIVER = os.path.join(OUT, "bin", "tlfea_driver")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source (always; no staleness check), link the shared library and the
    drivers."""
    os.makedirs(OUT, exist_ok=True)
    hip, cpp = sources()
    srcs = hip + cpp
    jobs = jobs or min(8, max(1, len(srcs)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    # command-line drivers over the drop-in C++ API, linked against the library
    bindir = os.path.join(OUT, "bin")
    os.makedirs(bindir, exist_ok=True)
    for src in sorted(glob.glob(os.path.join(CSRC, "drivers", "*.cpp"))):
        exe = os.path.join(bindir, os.path.splitext(os.path.basename(src))[0])
        cmd = [HIPCC, "-x", "hip", *COMMON_FLAGS, src, "-o", exe + ".tmp", "-L", OUT, "-ltlfea_hip",
               "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed for {src}:\n{r.stderr[-6000:]}")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
