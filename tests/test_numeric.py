"""Intrinsic semantics of fortran-to-cuda_amd/csrc/numeric.hpp (compiler/runtime/numeric.hpp:11-112).

The division samples restate the construction of compiler/tests/test_exact_division.py:50-66 (special bit
patterns, every power-of-two exponent with both signs, xorshift fill); the reference's expected output for
them is "PASS 4 32768" and "PASS 8 147456" (test_exact_division.py:96). The boundary cases are those of
compiler/tests/test_intrinsics.py:145-189, plus the signed-zero and NaN orderings of the min/max folds. The
host build runs here; the device build runs the same samples through ftc_divide_probe / ftc_intrinsic_probe
(GPU tests below).
"""
import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "fortran-to-cuda_amd" / "csrc"

HOST_PROGRAM = r"""
#include "numeric.hpp"
#include <cassert>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
using namespace ftc::numeric;
template <class T, class U, int M, int E> void check() {
    constexpr U sign = U(1) << (sizeof(U) * 8 - 1), exponent = U((1 << E) - 1) << M;
    const U special[] = {0, sign, 1, sign | 1, (U(1) << M) - 1, U(1) << M, exponent - 1, sign | (exponent - 1),
                         exponent, sign | exponent, exponent | 1, exponent | (U(1) << (M - 1)), exponent | 12345,
                         sign | exponent | 54321, U((1 << (E - 1)) - 1) << M, sign | (U((1 << (E - 1)) - 1) << M)};
    const int powers = (1 << E) * 32, n = powers + 8192;
    unsigned long long state = 0x123456789abcdef0ULL;
    auto rnd = [&]() { state ^= state << 13; state ^= state >> 7; state ^= state << 17; return U(state); };
    int compared = 0;
    for (int i = 0; i < n; ++i) {
        U av = i % 32 < 16 ? special[i % 16] : rnd();
        U bv = i < powers ? (U(i / 32) << M) | ((i & 1) ? sign : 0) : i % 32 < 16 ? special[i % 16] : rnd();
        T a, b;
        std::memcpy(&a, &av, sizeof(T));
        std::memcpy(&b, &bv, sizeof(T));
        volatile T va = a, vb = b;
        const T r1 = va / vb, c1 = divide(a, b);
        const T r2 = T(i - 41) / vb, c2 = divide(i - 41, b);
        if (std::memcmp(&r1, &c1, sizeof(T)) || std::memcmp(&r2, &c2, sizeof(T))) {
            std::printf("mismatch %zu %d\n", sizeof(T), i);
            std::exit(2);
        }
        compared += 2;
    }
    std::printf("PASS %zu %d\n", sizeof(T), compared);
}
int main() {
    volatile int minimum_int = INT_MIN;
    const int imin = minimum_int;
    assert(mod(imin, -1) == 0);
    assert(modulo(imin, -1) == 0);
    assert(sign(imin, -1) == imin);
    assert(dim(imin, 1) == 0);
    assert(nint(std::nextafter(0.5, 0.0)) == 0);
    assert(nint(std::nextafter(-0.5, 0.0)) == 0);
    assert(nint(1.5) == 2 && nint(-1.5) == -2);
    assert(std::signbit(sign(1.0, -0.0)));
    assert(std::signbit(modulo(0.0, -1.0)));
    assert(modulo(-7, 3) == 2 && modulo(7, -3) == -2 && mod(-7, 3) == -1);
    const double qnan = std::nan("");
    assert(std::isnan(minimum(1.0, qnan)) && minimum(qnan, 1.0) == 1.0);
    assert(std::signbit(minimum(0.0, -0.0)) && !std::signbit(minimum(-0.0, 0.0)));
    volatile double zero = 0.0;
    const double nz = -zero, rnan = nz / nz;
    assert(!std::signbit(maximum(nz, 0.0)) && std::signbit(maximum(0.0, nz)));
    assert(!std::signbit(minimum_constant_first(0.0, nz)) && std::signbit(absolute(nz)));
    assert(std::isnan(minimum(1.0, rnan)) && minimum(rnan, 1.0) == 1.0 && minimum_constant_first(1.0, rnan) == 1.0);
    int a = 0, p = 0, mask = 0;      // each argument evaluated once
    assert(modulo(++a, ++p) == 0 && a == 1 && p == 1);
    assert(merge(++a, ++p, ++mask) == 2 && a == 2 && p == 2 && mask == 1);
    check<float, unsigned, 23, 8>();
    check<double, unsigned long long, 52, 11>();
}
"""


def test_host_build_matches_the_reference_division_and_intrinsic_cases(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "numeric_host.cpp"
    src.write_text(HOST_PROGRAM)
    exe = tmp_path / "numeric_host"
    subprocess.run([gxx, "-std=c++17", "-O2", "-ffp-contract=off", "-fsanitize=undefined",
                    "-fno-sanitize-recover=undefined", "-I", str(CSRC), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert [line for line in out if line] == ["PASS 4 32768", "PASS 8 147456"]


def division_samples(width):
    """The (a, b) bit patterns of test_exact_division.py:50-66 as numpy arrays of the given width."""
    if width == 4:
        U, M, E, dtype = np.uint32, 23, 8, np.float32
    else:
        U, M, E, dtype = np.uint64, 52, 11, np.float64
    bits = 32 if width == 4 else 64
    mask = (1 << bits) - 1
    sign = 1 << (bits - 1)
    exponent = ((1 << E) - 1) << M
    special = [0, sign, 1, sign | 1, (1 << M) - 1, 1 << M, exponent - 1, sign | (exponent - 1), exponent,
               sign | exponent, exponent | 1, exponent | (1 << (M - 1)), exponent | 12345, sign | exponent | 54321,
               ((1 << (E - 1)) - 1) << M, sign | (((1 << (E - 1)) - 1) << M)]
    powers = (1 << E) * 32
    n = powers + 8192
    state = 0x123456789ABCDEF0
    m64 = (1 << 64) - 1

    def rnd():
        nonlocal state
        state ^= (state << 13) & m64
        state ^= state >> 7
        state ^= (state << 17) & m64
        return state & mask

    a = np.empty(n, dtype=U)
    b = np.empty(n, dtype=U)
    for i in range(n):
        a[i] = special[i % 16] if i % 32 < 16 else rnd()
        if i < powers:
            b[i] = ((i // 32) << M) | (sign if i & 1 else 0)
        else:
            b[i] = special[i % 16] if i % 32 < 16 else rnd()
    return a.view(dtype), b.view(dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("width,expected", [(4, 32768), (8, 147456)])
def test_device_division_is_bit_identical_to_plain_division(hip, width, expected):
    a, b = division_samples(width)
    n = len(a)
    ref = np.zeros(2 * n, dtype=a.dtype)
    cand = np.zeros(2 * n, dtype=a.dtype)
    ptr = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    assert hip.ftc_divide_probe(width, ptr(a), ptr(b), ptr(ref), ptr(cand), n) == 0
    assert 2 * n == expected
    u = np.uint32 if width == 4 else np.uint64
    assert np.array_equal(ref.view(u), cand.view(u))
    # The plain device division is IEEE division: it matches the host's on every finite result, and is a
    # NaN exactly where the host's is.
    with np.errstate(all="ignore"):
        host = np.concatenate([a / b, (np.arange(n) - 41).astype(a.dtype) / b])
    nan = np.isnan(host)
    assert np.array_equal(np.isnan(ref), nan)
    assert np.array_equal(ref[~nan].view(u), host[~nan].view(u))


@pytest.mark.gpu
def test_device_intrinsic_boundary_cases(hip):
    failures = ctypes.c_int(-1)
    assert hip.ftc_intrinsic_probe(ctypes.byref(failures)) == 0
    assert failures.value == 0, bin(failures.value)
