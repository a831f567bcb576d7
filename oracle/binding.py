"""ORACLE — test infrastructure only.

ctypes bindings of the oracle's C restatement (oracle/liboracle.so) and of the reference's own Fortran
built into oracle/_ref/. Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline load this
module, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_void_p
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "liboracle.so"
REF = ORACLE_DIR / "_ref"

_V = c_void_p
_D = c_double
_I = c_int

SIGNATURES = {
    "orc_cdu": [_V, _V, _V, _V, _D, _D, _D, _I, _I, _I, _I, _I, _I],
    "orc_cdv": [_V, _V, _V, _V, _D, _D, _D, _I, _I, _I, _I, _I, _I],
    "orc_cdw": [_V, _V, _V, _V, _D, _D, _D, _I, _I, _I, _I, _I, _I],
    "orc_scalar_adv": [_V, _V, _V, _V, _V, _D, _D, _D, _I, _I, _I, _I, _I, _I],
    "orc_sgs_coefficients": [_D, _D, _D, _D, _V],
    "orc_sgs_nu": [_V, _V, _V, _V, _V, _I, _I, _I, _I, _I, _I],
    "orc_sgs_diffusion": [_V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _I, _I, _I, _I, _I, _I],
    "orc_halo_les": [_V, _I, _I, _I, _I, _I, _I, _I],
    "orc_box_periodic": [_V, _I, _I, _I],
    "orc_box_diffuse_store": [_V, _V, _I, _I, _I, _D, _D, _D],
    "orc_les_tendency": [_V, _V, _V, _V, _V, _V, _V, _V, _V, _D, _D, _D, _V, _I, _I, _I],
    "orc_les_step": [_V, _V, _V, _V, _V, _V, _V, _V, _D, _D, _D, _D, _V, _I, _I, _I],
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: run `make -C oracle`")
        _lib = ctypes.CDLL(str(LIB))
        for name, args in SIGNATURES.items():
            fn = getattr(_lib, name, None)
            if fn is not None:
                fn.restype = None
                fn.argtypes = args
    return _lib


def divergence(U, V, W, dx, dy, dz, dt) -> np.ndarray:
    nx, ny, nz = (d - 2 for d in U.shape)
    rhs = np.zeros((nx, ny, nz), order="F")
    fn = lib().orc_divergence
    fn.restype = None
    fn.argtypes = [_V, _V, _V, _V, _D, _D, _D, _D, _I, _I, _I]
    fn(ptr(rhs), ptr(U), ptr(V), ptr(W), dx, dy, dz, dt, nx, ny, nz)
    return rhs


def projection(U, V, W, phi, dx, dy, dz, dt) -> None:
    """In place: U -= dt grad(phi) on the interior faces, then the halos (oracle/step.c)."""
    nx, ny, nz = (d - 2 for d in U.shape)
    phi = np.asfortranarray(phi)
    fn = lib().orc_projection
    fn.restype = None
    fn.argtypes = [_V, _V, _V, _V, _D, _D, _D, _D, _I, _I, _I]
    fn(ptr(U), ptr(V), ptr(W), ptr(phi), dx, dy, dz, dt, nx, ny, nz)


def projected_step(U, V, W, S, dx, dy, dz, dt, cs):
    """One step of the ensemble: explicit LES step, divergence, FP64 pressure (oracle/poisson.py),
    projection. Returns (U, V, W, S, phi)."""
    from oracle import poisson

    Un, Vn, Wn, Sn = les_step(U, V, W, S, dx, dy, dz, dt, cs)
    rhs = divergence(Un, Vn, Wn, dx, dy, dz, dt)
    phi = np.asfortranarray(poisson.solve(rhs, (dx, dy, dz)))
    projection(Un, Vn, Wn, phi, dx, dy, dz, dt)
    return Un, Vn, Wn, Sn, phi


def fixed_order_reduce(x: np.ndarray, lo, hi, op: int, initial: float) -> float:
    """The fixed-order reduction of the box [lo, hi) of a rank-3 Fortran array (oracle/reduce.c)."""
    l = lib()
    fn = l.orc_fixed_order_reduce
    fn.restype = c_double
    fn.argtypes = [_V, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                   ctypes.c_longlong, _I, _D]
    e1, e2 = x.shape[0], x.shape[1]
    base = x.ctypes.data + 8 * (lo[0] + lo[1] * e1 + lo[2] * e1 * e2)
    return fn(base, hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], e1, e1 * e2, op, initial)


def box_step(a: np.ndarray, b: np.ndarray, dx: float, dt: float, nu: float) -> None:
    """One step of the box application without sources, in place: the two-layer periodic halo of `a`
    (boundaries.f90:6-41), then b = a + dt*nu*laplacian(a) and the interior of a := b (stepper.f90:8-37).
    Arrays are a(-1:nx+2, -1:ny+2, -1:nz+2), Fortran order."""
    nx, ny, nz = (d - 4 for d in a.shape)
    assert a.flags.f_contiguous and b.flags.f_contiguous and a.shape == b.shape
    lib().orc_box_periodic(a.ctypes.data, nx, ny, nz)
    lib().orc_box_diffuse_store(a.ctypes.data, b.ctypes.data, nx, ny, nz, dx, dt, nu)


def ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def sgs_coefficients(dx, dy, dz, cs) -> np.ndarray:
    out = np.zeros(10)
    lib().orc_sgs_coefficients(dx, dy, dz, cs, ptr(out))
    return out


def les_tendency(U, V, W, S, dx, dy, dz, cs=None):
    """Tendencies (TU, TV, TW, TS) of the LES fields; cs None means advection alone."""
    nx, ny, nz = (d - 2 for d in U.shape)
    outs = [np.zeros_like(U, order="F") for _ in range(5)]
    coef = None if cs is None else sgs_coefficients(dx, dy, dz, cs)
    lib().orc_les_tendency(*(ptr(o) for o in outs), ptr(U), ptr(V), ptr(W), ptr(S), dx, dy, dz, ptr(coef),
                           nx, ny, nz)
    return outs[:4]


def les_step(U, V, W, S, dx, dy, dz, dt, cs=None):
    nx, ny, nz = (d - 2 for d in U.shape)
    outs = [np.zeros_like(U, order="F") for _ in range(4)]
    coef = None if cs is None else sgs_coefficients(dx, dy, dz, cs)
    lib().orc_les_step(*(ptr(o) for o in outs), ptr(U), ptr(V), ptr(W), ptr(S), dx, dy, dz, dt, ptr(coef),
                       nx, ny, nz)
    return outs


def component(name: str, U, V, W, dx, dy, dz, n1, n2, n3, out=None):
    """One reference advection component (cdu, cdv, cdw) over loop extents n1, n2, n3."""
    out = np.zeros_like(U, order="F") if out is None else out
    e1, e2, e3 = U.shape
    getattr(lib(), "orc_" + name)(ptr(out), ptr(U), ptr(V), ptr(W), dx, dy, dz, n1, n2, n3, e1, e2, e3)
    return out


def sgs_nu(U, V, W, dx, dy, dz, cs):
    nx, ny, nz = (d - 2 for d in U.shape)
    NU = np.zeros_like(U, order="F")
    coef = sgs_coefficients(dx, dy, dz, cs)
    lib().orc_sgs_nu(ptr(NU), ptr(U), ptr(V), ptr(W), ptr(coef), nx, ny, nz, *U.shape)
    return NU


def halo_les(F, kind: int):
    nx, ny, nz = (d - 2 for d in F.shape)
    lib().orc_halo_les(ptr(F), kind, nx, ny, nz, *F.shape)


# ---------------------------------------------------------------------------------- reference builds
_refs: dict[str, ctypes.CDLL] = {}


def reference_component(name: str):
    """The reference's own CDU/CDV/CDW built by `make -C oracle ref`, or None if it was not built."""
    path = REF / f"libref_{name}.so"
    if not path.exists():
        return None
    if name not in _refs:
        l = ctypes.CDLL(str(path))
        fn = getattr(l, "ref_" + name)
        fn.restype = None
        fn.argtypes = [_V, _V, _V, _V, _D, _D, _D, _I, _I, _I, _I, _I, _I]
        _refs[name] = l
    fn = getattr(_refs[name], "ref_" + name)

    def call(U, V, W, dx, dy, dz, n1, n2, n3, out=None):
        out = np.zeros_like(U, order="F") if out is None else out
        fn(ptr(out), ptr(U), ptr(V), ptr(W), dx, dy, dz, n1, n2, n3, *U.shape)
        return out

    return call
