"""Builds libhg.so (the HIP engine, gfx950) and the C++ host tools in-tree.

Run as ``python -m libhybridgrains_amd.build`` or through
``__graft_entry__.build()``.  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
LIB = PKG / "libhg.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off keeps every product and sum rounded as written, so the
# FP64 path follows the same operation order as the CPU restatement.
# Packed FP32 instructions (v_pk_*_f32) are disabled: some of their forms have
# given wrong results on the target machines without any error.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]

HIP_SOURCES = ["hg_ctx.hip", "hg_sort.hip", "hg_mpm.hip", "hg_dem.hip", "hg_hybrid.hip", "hg_server.hip",
               "hg_snapshot.hip", "hg_observe.hip"]

HOST_LIB = PKG / "libhg_host.so"
HOST_SOURCES = ["hg_xml.cpp", "hg_scene.cpp", "hg_sim.cpp", "hg_resample.cpp", "hg_host_capi.cpp"]
CLIS = ["mpmgrains2d_cli", "rigidbody2d_cli", "hybridgrains2d_cli"]
CXX = os.environ.get("CXX", "g++")
ROCM_LIB = "/opt/rocm/lib"
# the host tools run the zone resampling's independent per-trial and
# per-point loops on OpenMP threads
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-fopenmp", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_lib(jobs: int = 8) -> Path:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError(f"hipcc failed on {src}")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_host() -> Path:
    """libhg_host.so (scene readers, reference-named stepping API) and the
    command-line drivers under bin/; linked against libhg.so."""
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run([CXX, *CXX_FLAGS, "-shared", "-o", str(tmp), *[str(HOST / s) for s in HOST_SOURCES],
          f"-L{PKG}", "-lhg", f"-L{ROCM_LIB}", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM_LIB}"])
    os.replace(tmp, HOST_LIB)
    bindir = PKG / "bin"
    bindir.mkdir(exist_ok=True)
    for name in CLIS:
        out = bindir / name
        _run([CXX, *CXX_FLAGS, "-o", str(out) + ".tmp", str(HOST / f"{name}.cpp"), f"-L{PKG}",
              "-lhg_host", "-lhg", f"-L{ROCM_LIB}", "-Wl,-rpath,$ORIGIN/..",
              f"-Wl,-rpath,{ROCM_LIB}"])
        os.replace(str(out) + ".tmp", out)
    return HOST_LIB


def main() -> int:
    build_lib()
    build_host()
    return 0


if __name__ == "__main__":
    sys.exit(main())
