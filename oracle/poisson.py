"""ORACLE — test infrastructure only.

FP64 direct solve of the pressure Poisson problem on the CPU, written independently of the device
method: complex FFTs (periodic) or orthonormal DCT-II (Neumann, staggered) along x and y, then one
tridiagonal system per (kx, ky) mode along z (Neumann, staggered) solved by the Thomas algorithm,
vectorised over modes; the singular constant mode is solved by a minimum-norm least-squares solve,
which is the zero-mean solution.

The reference's solver is the poisson-solver submodule (VkFFT, not in the reference; call sites
benchmarks/elmm/sources/gpu_pressure.cpp:74,107-108,131-132). Its discretisation is stated there: FD2,
periodic (0) or NeumannStag (4) boundary pairs (gpu_pressure.cpp:51-69). Parity with that submodule is
unpinned (its code is absent); this oracle is pinned instead by tests/test_poisson_oracle.py: the discrete
eigenvectors (a manufactured solution whose answer is exact), an independent stencil residual, and a
sparse direct solve (scipy.sparse.linalg.spsolve) of the same FD2 system with the mean fixed.
"""
from __future__ import annotations

import numpy as np
import scipy.fft


def laplacian(phi: np.ndarray, h, bc=("periodic", "periodic", "neumann")) -> np.ndarray:
    """The FD2 7-point operator on interior cells; neighbours wrap (periodic) or mirror (Neumann)."""
    out = np.zeros_like(phi)
    for axis, (hd, kind) in enumerate(zip(h, bc)):
        if kind == "periodic":
            lo = np.roll(phi, 1, axis=axis)
            hi = np.roll(phi, -1, axis=axis)
        else:
            pad = np.concatenate([np.take(phi, [0], axis=axis), phi, np.take(phi, [-1], axis=axis)], axis=axis)
            n = phi.shape[axis]
            lo = np.take(pad, range(0, n), axis=axis)
            hi = np.take(pad, range(2, n + 2), axis=axis)
        out += ((lo + hi) - 2.0 * phi) / (hd * hd)
    return out


def _eigen(n: int, h: float, kind: str) -> np.ndarray:
    m = np.arange(n)
    if kind == "periodic":
        return (2.0 * np.cos(2.0 * np.pi * m / n) - 2.0) / (h * h)
    return (2.0 * np.cos(np.pi * m / n) - 2.0) / (h * h)


def _forward(a: np.ndarray, axis: int, kind: str) -> np.ndarray:
    if kind == "periodic":
        return np.fft.fft(a, axis=axis)
    return scipy.fft.dct(a, type=2, axis=axis, norm="ortho")


def _inverse(a: np.ndarray, axis: int, kind: str) -> np.ndarray:
    if kind == "periodic":
        return np.fft.ifft(a, axis=axis)
    return scipy.fft.idct(a, type=2, axis=axis, norm="ortho")


def solve(rhs: np.ndarray, h, bc=("periodic", "periodic")) -> np.ndarray:
    """phi (zero mean) with laplacian(phi) = rhs on the interior; rhs of shape (nx, ny, nz), zero mean."""
    nx, ny, nz = rhs.shape
    hx, hy, hz = h
    B = _forward(_forward(rhs.astype(np.complex128), 0, bc[0]), 1, bc[1])
    lam = _eigen(nx, hx, bc[0])[:, None] + _eigen(ny, hy, bc[1])[None, :]
    r = 1.0 / (hz * hz)
    # Thomas along z for every mode at once.
    diag = np.empty((nz,) + lam.shape)
    for k in range(nz):
        diag[k] = -2.0 * r + lam + (r if k == 0 else 0.0) + (r if k == nz - 1 else 0.0)
    zero = lam == 0.0
    # The singular column (lam == 0) is replaced below; until then its diagonal is shifted by -2r, which
    # makes the column strictly diagonally dominant at every grid spacing, so no pivot is zero.
    safe = np.where(zero, -2.0 * r, 0.0)
    cp = np.zeros((nz,) + lam.shape)
    dp = np.zeros((nz,) + lam.shape, dtype=np.complex128)
    denom = diag[0] + safe
    cp[0] = r / denom
    dp[0] = B[:, :, 0] / denom
    for k in range(1, nz):
        denom = diag[k] - r * cp[k - 1] + safe
        cp[k] = r / denom
        dp[k] = (B[:, :, k] - r * dp[k - 1]) / denom
    psi = np.zeros_like(dp)
    psi[-1] = dp[-1]
    for k in range(nz - 2, -1, -1):
        psi[k] = dp[k] - cp[k] * psi[k + 1]
    psi = np.moveaxis(psi, 0, 2)
    # The constant mode: minimum-norm least squares of the singular z operator.
    T = np.zeros((nz, nz))
    for k in range(nz):
        T[k, k] = -2.0 * r + (r if k == 0 else 0.0) + (r if k == nz - 1 else 0.0)
        if k > 0:
            T[k, k - 1] = r
        if k < nz - 1:
            T[k, k + 1] = r
    for p, q in zip(*np.nonzero(zero)):
        col = B[p, q, :]
        re = np.linalg.lstsq(T, col.real, rcond=None)[0]
        im = np.linalg.lstsq(T, col.imag, rcond=None)[0]
        psi[p, q, :] = re + 1j * im
    phi = _inverse(_inverse(psi, 1, bc[1]), 0, bc[0])
    return np.ascontiguousarray(phi.real)


def divergence_rhs(nx: int, ny: int, nz: int, h, seed: int) -> np.ndarray:
    """The configuration's right-hand side: the divergence of a seeded random staggered velocity field
    (periodic in x and y, w = 0 on the z walls), made zero-mean. Returns (nx, ny, nz)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, (nx, ny, nz))
    v = rng.uniform(-1.0, 1.0, (nx, ny, nz))
    w = rng.uniform(-1.0, 1.0, (nx, ny, nz + 1))
    w[:, :, 0] = 0.0
    w[:, :, -1] = 0.0
    div = ((u - np.roll(u, 1, axis=0)) / h[0] + (v - np.roll(v, 1, axis=1)) / h[1]
           + (w[:, :, 1:] - w[:, :, :-1]) / h[2])
    return div - div.mean()
