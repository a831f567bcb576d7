"""MI355X-native resident backend for Fortran LES loop nests.

The product is the C-ABI library ``libftc_hip.so`` built from ``csrc/`` (the fort_scope residency runtime
and the hand-written gfx950 kernels with their entries, declared in ``include/*.h`` at the repository
root). This Python package only loads it and offers thin callers for tests, the benchmark and Python
users; it has no compute path of its own and no fallback: when the library is missing, loading fails.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_uint64, c_void_p
from pathlib import Path

from . import scope

PACKAGE = Path(__file__).resolve().parent
HIP_LIBRARY = PACKAGE / "libftc_hip.so"
CPU_LIBRARY = PACKAGE / "libftc_cpu.so"

_DP = POINTER(c_double)
_IP = POINTER(c_int)
_H = c_uint64

# Entry signatures: (name, argtypes). Every entry returns an int fort_scope status.
ENTRY_SIGNATURES = {
    "cpp_fort_shared_momentumadvection_cdu": [c_uint64, c_int, _H, _H, _H, _H, _DP, _DP, _DP, _IP, _IP, _IP],
    "cpp_fort_shared_momentumadvection_cdv": [c_uint64, c_int, _H, _H, _H, _H, _DP, _DP, _DP, _IP, _IP, _IP],
    "cpp_fort_shared_momentumadvection_cdw": [c_uint64, c_int, _H, _H, _H, _H, _DP, _DP, _DP, _IP, _IP, _IP],
    "cpp_fort_shared_les_tendency": [c_uint64, c_int, _H, _H, _H, _H, _H, _H, _H, _H, _DP, _DP, _DP, _DP,
                                     _IP, _IP, _IP, _IP],
    "cpp_fort_shared_les_step": [c_uint64, c_int, _H, _H, _H, _H, _H, _H, _H, _H, _DP, _DP, _DP, _DP, _DP,
                                 _IP, _IP, _IP, _IP],
    "cpp_fort_shared_chain_contraction": [c_uint64, c_int, _H, _H, _H, _H, _IP, _IP],
    "cpp_fort_shared_pressure_divergence": [c_uint64, c_int, _H, _H, _H, _H, _DP, _DP, _DP, _DP, _IP, _IP, _IP],
    "cpp_fort_shared_pressure_projection": [c_uint64, c_int, _H, _H, _H, _H, _DP, _DP, _DP, _DP, _IP, _IP, _IP],
    "cpp_fort_shared_pressure_fix_reference": [c_uint64, c_int, _H, _DP, _IP, _IP, _IP],
    "cpp_fort_shared_fixed_order_reduction": [c_uint64, c_int, _H, _IP, POINTER(ctypes.c_longlong),
                                              POINTER(ctypes.c_longlong), _DP],
    "cpp_fort_shared_fixed_order_reduction_device": [c_uint64, c_int, _H, _H, _H, _IP, POINTER(ctypes.c_longlong),
                                                     POINTER(ctypes.c_longlong), _DP],
    "cpp_fort_shared_pressure_fix_reference_sum": [c_uint64, c_int, _H, _H, _IP, _IP, _IP],
    "cpp_fort_shared_pressure_projection_bc": [c_uint64, c_int, _H, _H, _H, _H, _DP, _DP, _DP, _DP, _IP, _IP, _IP,
                                               _IP, _IP],
}


class GemmDesc(ctypes.Structure):
    _fields_ = [("a", c_void_p), ("b", c_void_p), ("c", c_void_p), ("type_a", c_int), ("type_b", c_int),
                ("type_c", c_int), ("m", c_int), ("n", c_int), ("k", c_int), ("batch", c_int),
                ("lda", ctypes.c_longlong), ("ldb", ctypes.c_longlong), ("ldc", ctypes.c_longlong),
                ("stride_a", ctypes.c_longlong), ("stride_b", ctypes.c_longlong), ("stride_c", ctypes.c_longlong),
                ("alpha", c_double), ("beta", c_double), ("flags", c_int)]
HELPER_SIGNATURES = {
    "ftc_les_coefficients": (None, [c_double, c_double, c_double, c_double, _DP]),
    "ftc_event_create": (c_int, [POINTER(c_void_p)]),
    "ftc_event_record": (c_int, [c_void_p, c_void_p]),
    "ftc_event_elapsed_ms": (c_int, [c_void_p, c_void_p, POINTER(c_float)]),
    "ftc_event_destroy": (c_int, [c_void_p]),
    "ftc_device_count": (c_int, []),
    "ftc_copy_probe": (c_int, [ctypes.c_size_t, c_int, POINTER(c_float)]),
    "ftc_bandwidth_probe": (c_int, [c_int, ctypes.c_size_t, c_int, POINTER(c_float)]),
    "ftc_mfma_probe": (c_int, [c_int, c_int, c_int, POINTER(c_float), POINTER(ctypes.c_double)]),
    "ftc_tile_visit_probe_v1": (c_int, [c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                        POINTER(ctypes.c_uint)]),
    "ftc_gemm_device": (c_int, [POINTER(GemmDesc), c_int, c_int, c_void_p]),
    "ftc_gemm_tile_count": (c_int, []),
    "ftc_mfma_layout_probe": (c_int, [c_int, c_int, c_int, c_int, c_int, _DP]),
    "ftc_divide_probe": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "ftc_intrinsic_probe": (c_int, [_IP]),
    "ftc_fixed_order_reduction_static_v1": (c_int, [c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                                    ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                                    c_int, c_double, c_void_p, c_void_p]),
}

_loaded: dict[str, ctypes.CDLL] = {}


def load(path: Path | str | None = None) -> ctypes.CDLL:
    """Loads libftc_hip.so (the device library). Raises if it has not been built."""
    path = Path(path) if path else HIP_LIBRARY
    key = str(path)
    if key in _loaded:
        return _loaded[key]
    if not path.exists():
        raise RuntimeError(f"{path} is missing: run `python fortran-to-cuda_amd/build.py` (or "
                           "__graft_entry__.build()) first; this package has no non-HIP path")
    lib = ctypes.CDLL(str(path), mode=ctypes.RTLD_GLOBAL)
    scope.bind_runtime(lib)
    for name, args in ENTRY_SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = c_int
            fn.argtypes = args
    for name, (res, args) in HELPER_SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args
    _loaded[key] = lib
    return lib


def load_cpu_backend() -> ctypes.CDLL:
    """Loads the CPU reference backend of the runtime (no kernels), for tests without a GPU."""
    if not CPU_LIBRARY.exists():
        raise RuntimeError(f"{CPU_LIBRARY} is missing: build the package first")
    key = str(CPU_LIBRARY)
    if key not in _loaded:
        _loaded[key] = scope.bind_runtime(ctypes.CDLL(key))
    return _loaded[key]


def dp(value: float) -> ctypes.Array:
    return (c_double * 1)(value)


def ip(value: int) -> ctypes.Array:
    return (c_int * 1)(value)
