"""ctypes bindings of the batched ensemble (include/ftc_ensemble.h)."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

FP64_PER_MEMBER, FP64_BATCHED, MIXED_BATCHED = 0, 1, 2


class Info(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("mode", c_uint32), ("members", c_int32), ("solves", c_int32),
                ("les_kchunk", c_int32), ("gemm_tile", c_int32), ("blocks_per_cu", c_int32), ("les_policy", c_int32),
                ("launches_per_step", c_int64), ("cells", c_int64), ("algorithmic_bytes", c_double),
                ("device_bytes", c_uint64)]


_DP = POINTER(c_double)
SIGNATURES = {
    "ftc_ensemble_create_v1": (c_int, [c_int, POINTER(c_int32), c_double, c_double, c_double, c_double, c_double,
                                       c_int, POINTER(c_void_p)]),
    "ftc_ensemble_upload_v1": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ftc_ensemble_download_v1": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ftc_ensemble_download_pressure_v1": (c_int, [c_void_p, c_int, c_void_p]),
    "ftc_ensemble_set_solves_v1": (c_int, [c_void_p, c_int]),
    "ftc_ensemble_set_projection_v1": (c_int, [c_void_p, c_int]),
    "ftc_ensemble_set_les_policy_v1": (c_int, [c_void_p, c_int]),
    "ftc_ensemble_autotune_v1": (c_int, [c_void_p, c_char_p]),
    "ftc_ensemble_step_v1": (c_int, [c_void_p, c_int]),
    "ftc_ensemble_synchronize_v1": (c_int, [c_void_p]),
    "ftc_ensemble_stream_v1": (c_int, [c_void_p, POINTER(c_void_p)]),
    "ftc_ensemble_get_info_v1": (c_int, [c_void_p, POINTER(Info)]),
    "ftc_ensemble_destroy_v1": (c_int, [POINTER(c_void_p)]),
    "ftc_ensemble_error_v1": (c_char_p, []),
}


def bind(lib):
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class Ensemble:
    def __init__(self, lib, sizes, d=(10.0, 10.0, 10.0), dt=0.2, cs=0.17, mode=MIXED_BATCHED):
        self.lib = bind(lib)
        self.sizes = [tuple(int(x) for x in s) for s in sizes]
        flat = (c_int32 * (3 * len(sizes)))(*[x for s in self.sizes for x in s])
        self.handle = c_void_p()
        self._check(lib.ftc_ensemble_create_v1(len(sizes), flat, d[0], d[1], d[2], dt, cs, mode,
                                               ctypes.byref(self.handle)))

    def _check(self, status):
        if status:
            raise RuntimeError(f"ensemble error {status}: {self.lib.ftc_ensemble_error_v1().decode()}")

    def upload(self, member, U, V, W, S):
        arrays = [np.asfortranarray(a, dtype=np.float64) for a in (U, V, W, S)]
        self._check(self.lib.ftc_ensemble_upload_v1(self.handle, member, *(a.ctypes.data for a in arrays)))

    def download(self, member):
        nx, ny, nz = self.sizes[member]
        out = [np.zeros((nx + 2, ny + 2, nz + 2), order="F") for _ in range(4)]
        self._check(self.lib.ftc_ensemble_download_v1(self.handle, member, *(a.ctypes.data for a in out)))
        return out

    def pressure(self, member):
        nx, ny, nz = self.sizes[member]
        phi = np.zeros((nx, ny, nz), order="F")
        self._check(self.lib.ftc_ensemble_download_pressure_v1(self.handle, member, phi.ctypes.data))
        return phi

    def solves(self, n):
        self._check(self.lib.ftc_ensemble_set_solves_v1(self.handle, n))

    def projection(self, on: bool):
        self._check(self.lib.ftc_ensemble_set_projection_v1(self.handle, 1 if on else 0))

    def les_policy(self, policy: int):
        self._check(self.lib.ftc_ensemble_set_les_policy_v1(self.handle, policy))

    def autotune(self, cache=None):
        self._check(self.lib.ftc_ensemble_autotune_v1(self.handle, cache.encode() if cache else None))

    def step(self, n=1):
        self._check(self.lib.ftc_ensemble_step_v1(self.handle, n))

    def synchronize(self):
        self._check(self.lib.ftc_ensemble_synchronize_v1(self.handle))

    def stream(self):
        s = c_void_p()
        self._check(self.lib.ftc_ensemble_stream_v1(self.handle, ctypes.byref(s)))
        return s

    def info(self) -> Info:
        i = Info()
        self._check(self.lib.ftc_ensemble_get_info_v1(self.handle, ctypes.byref(i)))
        return i

    def destroy(self):
        if self.handle:
            self._check(self.lib.ftc_ensemble_destroy_v1(ctypes.byref(self.handle)))
