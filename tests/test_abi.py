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
