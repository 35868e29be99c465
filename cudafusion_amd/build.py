"""Build recipe of the gfx950 library ``libnrf.so``.

Compiles every HIP source under ``cudafusion_amd/csrc`` with hipcc for gfx950
into one shared library next to this file.  Device code is built with
``-ffp-contract=off`` (no FMA fusion, so FP32 results follow the written
operation order) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libnrf.so")

HIP_SOURCES = ["common.hip", "integrate.hip", "mc.hip", "warp.hip", "solver.hip", "pipeline.hip", "precision.hip",
               "learned.hip", "association.hip", "raster.hip", "track.hip",
               "raycast.hip", "graph.hip"]
CXX_SOURCES = ["dropin.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile all sources (always, even if a library exists) and link."""
    hipcc = _hipcc()
    os.makedirs(OBJDIR, exist_ok=True)
    inc = ["-I", INCLUDE, "-I", CSRC]
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        objs.append(obj)
        cmds.append([hipcc, *HIPCC_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJDIR, src + ".o")
        objs.append(obj)
        cmds.append([hipcc, "-x", "hip", *HIPCC_FLAGS, *inc, "-c", path, "-o", obj])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd))
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp,
          "-Wl,-rpath,/opt/rocm/lib", "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    # C++ program that exercises the drop-in class API (run by a GPU test)
    test_src = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
    if os.path.exists(test_src):
        exe = os.path.join(ROOT, "tests", "cpp", "test_dropin")
        _run([hipcc, "-O2", "-std=c++17", "-I", INCLUDE, test_src, "-o", exe + ".tmp", "-L", PKG, "-lnrf",
              "-Wl,-rpath,$ORIGIN/../../cudafusion_amd", "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
