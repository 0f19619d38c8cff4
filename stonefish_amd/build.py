"""Builds the native libraries in-tree.

``libsfhydro.so`` (the HIP kernels and the C ABI of ``include/sfhydro.h``) is
compiled for gfx950 with hipcc; it lands in ``stonefish_amd/_lib/`` so that it
travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled: hipcc may form them from scalar float
# code, and all HIP code here is built without them.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wno-unused-value",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIBS = {
    "libsfhydro.so": ["sfh_hydro.hip", "sfh_hull.cpp", "sfh_waves.hip"],
    "libsfmlp.so": ["sfh_mlp.hip"],
}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_hip(verbose: bool = False) -> list[Path]:
    """Compiles every HIP library for gfx950; returns the built paths."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    LIBDIR.mkdir(exist_ok=True)
    out = []
    for lib, srcs in LIBS.items():
        paths = [CSRC / s for s in srcs]
        if not all(p.exists() for p in paths):
            continue
        target = LIBDIR / lib
        tmp = LIBDIR / (lib + ".tmp")
        cmd = [HIPCC, *HIP_FLAGS, "-o", str(tmp), *map(str, paths)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        os.replace(tmp, target)
        out.append(target)
    return out


def build_cpp_check(verbose: bool = False) -> Path:
    """Builds the C++ interface check program (tests/cpp) against libsfhydro.so."""
    target = LIBDIR / "sfh_cpp_check"
    cmd = ["g++", "-O2", "-std=c++17", "-o", str(target), str(ROOT / "tests" / "cpp" / "sfh_cpp_check.cpp"),
           f"-L{LIBDIR}", "-lsfhydro", f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    return target


def build_oracle(verbose: bool = False) -> None:
    """Builds the CPU oracle (test infrastructure) and, where the reference
    checkout exists, oracle/_ref."""
    cmd = ["make", "-s", "-C", str(ROOT / "oracle"), "-B", "all"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_cpp_check(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
