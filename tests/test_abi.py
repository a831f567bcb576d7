"""The C-ABI libraries load and export every function the headers in include/ declare (no GPU calls)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"
DECL = re.compile(r"^(?:int|void|uint32_t|const char \*)\s*\**\s*(\w+)\s*\(", re.MULTILINE)


def declared(header: str) -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", (INCLUDE / header).read_text(), flags=re.S)
    return DECL.findall(text)


def exported(path: Path) -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


def headers():
    return sorted(p.name for p in INCLUDE.glob("*.h"))


@pytest.mark.parametrize("header", headers())
def test_device_library_exports_every_declared_function(ftc, header):
    names = declared(header)
    assert names, header
    missing = sorted(set(names) - exported(ftc.HIP_LIBRARY))
    assert not missing, f"{header}: not exported by libftc_hip.so: {missing}"


def test_cpu_backend_exports_the_runtime_abi(ftc):
    missing = sorted(set(declared("fort_scope.h")) - exported(ftc.CPU_LIBRARY))
    assert not missing


def test_device_library_loads_without_a_device(ftc):
    lib = ftc.load()
    assert lib.fort_scope_abi_version() == 1
    assert isinstance(lib.ftc_device_count(), int)


def test_headers_compile_as_c_and_cxx(tmp_path):
    for header in headers():
        src = tmp_path / "probe.c"
        src.write_text(f'#include "{header}"\nint main(void) {{ return 0; }}\n')
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", str(INCLUDE), "-c", str(src), "-o",
                        str(tmp_path / "probe.o")], check=True)
        subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-x", "c++", "-I", str(INCLUDE), "-c",
                        str(src), "-o", str(tmp_path / "probe2.o")], check=True)


def test_contraction_descriptor_is_checked_on_the_host(ftc):
    # Rejections happen before any device call, so they run without a GPU; the descriptor's layout (with
    # its trailing flags word) matches ftc_gemm_desc in ftc_dense.h.
    import ctypes

    lib = ftc.load()
    assert ctypes.sizeof(ftc.GemmDesc) == 128
    cases = [(0, (1, 1, 1), 0, b"FP64 contraction takes plain double operands"),
             (0, (0, 0, 0), 1, b"FP64 contraction takes plain double operands"),
             (1, (1, 1, 1), 2, b"a transposed C needs a K-contiguous A"),
             (1, (1, 1, 1), 4, b"unknown contraction layout flags"),
             (1, (0, 1, 0), 1, b"unsupported K-contiguous operand types")]
    for precision, (ta, tb, tc), flags, message in cases:
        d = ftc.GemmDesc(4096, 8192, 12288, ta, tb, tc, 8, 8, 8, 1, 8, 8, 8, 0, 0, 0, 1.0, 0.0, flags)
        assert lib.ftc_gemm_device(ctypes.byref(d), precision, -1, None) == 1
        assert lib.fort_scope_error() == message
    short = ftc.GemmDesc(4096, 8192, 12288, 1, 1, 1, 8, 8, 16, 1, 8, 16, 8, 0, 0, 0, 1.0, 0.0, 1)
    assert lib.ftc_gemm_device(ctypes.byref(short), 1, -1, None) == 1  # K-contiguous lda below K
    assert lib.fort_scope_error() == b"leading dimension below the row count"


def test_fortran_interfaces_call_the_library(ftc, tmp_path):
    # The bind(C) interface modules of fortran-to-cuda_amd/fortran/ compile, link against libftc_hip.so and
    # reach the entries from Fortran: native mode declines (9) before any device work, and an unknown mode
    # is an argument error (1). No GPU is needed for either.
    import shutil

    fc = shutil.which("gfortran") or shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not Path(fc).exists() and not shutil.which(fc):
        pytest.skip("no Fortran compiler")
    srcdir = ftc.PACKAGE / "fortran"
    prog = tmp_path / "probe.f90"
    prog.write_text("""program probe
  use iso_c_binding
  use fort_shared_momentumadvection_cdu, only: cdu => run
  use fort_shared_fixed_order_reduction, only: reduce => run
  implicit none
  real(c_double) :: d = 10.0d0, v = 0.0d0
  integer(c_int) :: n = 16, op = 0
  integer(c_long_long) :: lo(3) = 0, hi(3) = 1
  print '(i0)', cdu(0_c_int64_t, 0_c_int, 1_c_int64_t, 2_c_int64_t, 3_c_int64_t, 4_c_int64_t, d, d, d, n, n, n)
  print '(i0)', reduce(0_c_int64_t, 7_c_int, 1_c_int64_t, op, lo, hi, v)
end program probe
""")
    flag = "-J" if "gfortran" in fc else "-module-dir"
    objs = []
    for src in sorted(srcdir.glob("*.f90")) + [prog]:
        obj = tmp_path / (src.stem + ".o")
        subprocess.run([fc, "-c", flag, str(tmp_path), str(src), "-o", str(obj)], check=True, cwd=tmp_path)
        objs.append(str(obj))
    exe = tmp_path / "probe"
    subprocess.run([fc, *objs, "-o", str(exe), str(ftc.HIP_LIBRARY), f"-Wl,-rpath,{ftc.PACKAGE}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["9", "1"]
