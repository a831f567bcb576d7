"""Builds the native libraries of the package in-tree.

- ``libftc_hip.so``: the fort_scope runtime (HIP backend) and every hand-written gfx950 kernel with its
  C-ABI entry, compiled by hipcc for ``--offload-arch=gfx950``;
- ``libftc_cpu.so``: the same runtime built by g++ with ``-DFTC_CPU_BACKEND``, whose "device" memory is
  host memory; it runs the coherence, eviction and queue tests on machines without a GPU;
- ``fortran/*.f90``: the Fortran interface modules, compiled (syntax and interface check) when a Fortran
  compiler is present.

The numerical contract of every device file is the reference's separate-arithmetic contract
(compiler/emission/numerical_contract.py:13-15): no contraction of a*b+c, no fast math, IEEE division
and square root.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
BUILD = PACKAGE / "_build"
HIP_LIBRARY = PACKAGE / "libftc_hip.so"
CPU_LIBRARY = PACKAGE / "libftc_cpu.so"

HIP_SOURCES = ("runtime.cpp", "les.hip", "poisson.hip", "gemm.hip", "reduce.hip", "ensemble.hip", "timing.hip",
               "box.hip", "relax.hip", "numeric.hip", "pressure_step.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-DFTC_CPU_BACKEND"]


def _run(command: list[str]) -> None:
    result = subprocess.run(command, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(line for line in result.stderr.splitlines() if noise not in line)
    if result.returncode != 0:
        raise RuntimeError(f"command failed ({result.returncode}): {' '.join(command)}\n{err}")


def _compile(source: str) -> Path:
    BUILD.mkdir(exist_ok=True)
    obj = BUILD / (source + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / source), "-o", str(obj)])
    return obj


def build_hip(jobs: int = 8) -> Path:
    """Compiles every device source for gfx950 and links libftc_hip.so."""
    sources = [s for s in HIP_SOURCES if (CSRC / s).exists()]
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = HIP_LIBRARY.with_suffix(".so.tmp")
    # -Bsymbolic: calls between the library's own exported functions (device_begin -> access_begin_v1)
    # bind inside it, also when another library exporting the same C ABI is loaded globally first.
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", str(tmp), *map(str, objects)])
    os.replace(tmp, HIP_LIBRARY)
    return HIP_LIBRARY


def build_cpu() -> Path:
    """Builds the CPU reference backend of the runtime."""
    cxx = shutil.which("g++") or "g++"
    tmp = CPU_LIBRARY.with_suffix(".so.tmp")
    _run([cxx, *CXX_FLAGS, "-shared", "-Wl,-Bsymbolic", "-o", str(tmp), str(CSRC / "runtime.cpp")])
    os.replace(tmp, CPU_LIBRARY)
    return CPU_LIBRARY


def build_fortran() -> None:
    """Compiles the Fortran interface modules (a syntax and interface check of the bind(C) side)."""
    fc = shutil.which("gfortran") or shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not Path(fc).exists() and not shutil.which(fc):
        return
    out = BUILD / "fortran"
    out.mkdir(parents=True, exist_ok=True)
    # fortran/app/ holds the domain module of applications built on the reference's Fortran host layer;
    # resident_app.py compiles it with that layer.
    sources = sorted((PACKAGE / "fortran").glob("*.f90"))
    for src in sources:
        _run([fc, "-c", "-O2", "-J" if "gfortran" in fc else "-module-dir", str(out), str(src),
              "-o", str(out / (src.stem + ".o"))])


def build_all(jobs: int = 8) -> None:
    build_cpu()
    build_hip(jobs)
    build_fortran()


if __name__ == "__main__":
    build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print("built", HIP_LIBRARY, CPU_LIBRARY)
