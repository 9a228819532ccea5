"""Build libgdp_hip.so in-tree for gfx950.

    python -m getdp_amd.build        (or __graft_entry__.build())

HIP sources are compiled with hipcc for gfx950 only; host-only C++ sources
with g++.  Packed FP32 instructions are disabled for all device code.  The
assembly and ILU-factorisation kernels are compiled without floating-point
contraction so that their arithmetic matches the reference's operation by
operation.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INC = ROOT / "include"
BUILD = ROOT / "_build"
LIB = PKG / "libgdp_hip.so"
ARCH = os.environ.get("GDP_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{INC}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             f"-I{INC}", f"-I{CSRC}"]

# (source, extra flags)
SOURCES = [
    ("runtime.hip", []),
    ("sell.hip", []),
    ("fe.hip", ["-ffp-contract=off"]),
    ("ilu.hip", ["-ffp-contract=off"]),
    ("krylov.hip", []),
    ("solve.hip", []),
    ("unet.hip", []),
    ("batch.hip", []),
    ("edges.hip", []),
    ("linalg_ops.hip", ["-ffp-contract=off"]),
    ("host_prep.cpp", []),
    ("ilut.cpp", ["-ffp-contract=off"]),
    ("mesh_io.cpp", []),
    ("res_io.cpp", []),
    ("solver_params.cpp", []),
]


def _compile(src: str, extra):
    s = CSRC / src
    o = BUILD / (src + ".o")
    if s.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", str(s), "-o", str(o)]
    else:
        cmd = ["g++", *CXX_FLAGS, *extra, "-c", str(s), "-o", str(o)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return o


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    present = [(s, f) for s, f in SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(present))) as ex:
        objs = list(ex.map(lambda sf: _compile(*sf), present))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", "-Wl,-z,defs", f"--offload-arch={ARCH}", *map(str, objs),
           "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
