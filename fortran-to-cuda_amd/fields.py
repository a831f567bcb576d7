"""Seeded synthetic LES inputs.

The generator never uses a compiler's random_number (whose stream differs between Fortran compilers): a
splitmix64 stream (Steele, Lea and Flood 2014) drives every draw, so the same seed gives the same bytes on
every machine with the same libm, and the arrays are handed unchanged to the CPU oracle and to the device.

Fields on the staggered grid of nx x ny x nz cells with one ghost layer, spacing dx, dy, dz:
  u = 5 + sum_{m=1..8} sin(2 pi (kx x/Lx + ky y/Ly) + phi_m)        (U at x faces)
  v =     sum_{m=1..8} sin(2 pi (kx x/Lx + ky y/Ly) + phi_m)        (V at y faces)
  w =     sum_{m=1..8} sin(2 pi (kx x/Lx + ky y/Ly) + phi_m) sin(pi kz z/Lz)   (W at z faces)
with integer kx, ky in [1, 4] (periodic in x and y), kz in [1, 3] (w vanishes on the walls), phases
uniform in [0, 2 pi); the scalar is uniform(0, 1) smoothed once by a 7-point mean (periodic in x, y,
zero-gradient in z). Halos are then filled as the LES step fills them.
"""
from __future__ import annotations

import numpy as np

SEED = 20261123
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """`count` consecutive outputs of splitmix64 started at `seed`, skipping `offset` outputs."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Doubles uniform in [0, 1) with 53 random bits each."""
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def halo_fill(F: np.ndarray, kind: str) -> None:
    """One ghost layer: periodic in x, then y; z walls by kind: 'tangential' (mirror with sign change),
    'normal' (zero on and beyond the walls), 'scalar' (zero gradient)."""
    nx, ny, nz = F.shape[0] - 2, F.shape[1] - 2, F.shape[2] - 2
    F[0, 1:ny + 1, 1:nz + 1] = F[nx, 1:ny + 1, 1:nz + 1]
    F[nx + 1, 1:ny + 1, 1:nz + 1] = F[1, 1:ny + 1, 1:nz + 1]
    F[:, 0, 1:nz + 1] = F[:, ny, 1:nz + 1]
    F[:, ny + 1, 1:nz + 1] = F[:, 1, 1:nz + 1]
    if kind == "tangential":
        F[:, :, 0] = -F[:, :, 1]
        F[:, :, nz + 1] = -F[:, :, nz]
    elif kind == "normal":
        F[:, :, 0] = 0.0
        F[:, :, nz] = 0.0
        F[:, :, nz + 1] = 0.0
    elif kind == "scalar":
        F[:, :, 0] = F[:, :, 1]
        F[:, :, nz + 1] = F[:, :, nz]
    else:
        raise ValueError(kind)


def les_fields(nx: int, ny: int, nz: int, dx: float = 10.0, dy: float = 10.0, dz: float = 10.0,
               seed: int = SEED, modes: int = 8, mean_wind: float = 5.0):
    """(U, V, W, S) as Fortran-ordered float64 arrays of shape (nx+2, ny+2, nz+2), halos filled."""
    draws = uniform(seed, 64 * modes + 8)
    pos = 0

    def take(n):
        nonlocal pos
        out = draws[pos:pos + n]
        pos += n
        return out

    lx, ly, lz = nx * dx, ny * dy, nz * dz
    i = np.arange(nx + 2, dtype=np.float64)
    j = np.arange(ny + 2, dtype=np.float64)
    k = np.arange(nz + 2, dtype=np.float64)
    # Fortran index i (1-based) of cell i spans [(i-2)dx, (i-1)dx]; with 0-based index n = i-1 the cell
    # centre is (n-0.5)dx and its upper x face is n*dx.
    xc, xf = (i - 0.5) * dx, i * dx
    yc, yf = (j - 0.5) * dy, j * dy
    zc, zf = (k - 0.5) * dz, k * dz

    def field(xs, ys, zprofile):
        out = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
        for _ in range(modes):
            kx = 1 + int(take(1)[0] * 4)
            ky = 1 + int(take(1)[0] * 4)
            kz = 1 + int(take(1)[0] * 3)
            phase = 2.0 * np.pi * take(1)[0]
            plane = np.sin(2.0 * np.pi * (kx * xs[:, None] / lx + ky * ys[None, :] / ly) + phase)
            out += plane[:, :, None] * zprofile(kz)[None, None, :]
        return out

    ones = lambda kz: np.ones(nz + 2)
    U = field(xf, yc, ones)
    U += mean_wind
    V = field(xc, yf, ones)
    W = field(xc, yc, lambda kz: np.sin(np.pi * kz * zf / lz))
    raw = uniform(seed ^ 0x5CA1A7, nx * ny * nz).reshape((nx, ny, nz), order="F")
    S = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    S[1:-1, 1:-1, 1:-1] = raw
    halo_fill(S, "scalar")
    c = S[1:-1, 1:-1, 1:-1]
    smooth = (c + S[:-2, 1:-1, 1:-1] + S[2:, 1:-1, 1:-1] + S[1:-1, :-2, 1:-1] + S[1:-1, 2:, 1:-1]
              + S[1:-1, 1:-1, :-2] + S[1:-1, 1:-1, 2:]) / 7.0
    S[1:-1, 1:-1, 1:-1] = smooth
    for F, kind in ((U, "tangential"), (V, "tangential"), (W, "normal"), (S, "scalar")):
        halo_fill(F, kind)
    return U, V, W, S


def divergence_rhs(nx: int, ny: int, nz: int, h, seed: int = SEED) -> np.ndarray:
    """A pressure right-hand side: the divergence of a seeded uniform(-1, 1) staggered velocity field
    (periodic in x and y, w = 0 on the z walls), minus its mean. Returns an (nx, ny, nz) Fortran array."""
    d = uniform(seed ^ 0xD1F, 3 * nx * ny * nz + nx * ny)
    u = (2.0 * d[:nx * ny * nz] - 1.0).reshape((nx, ny, nz), order="F")
    v = (2.0 * d[nx * ny * nz:2 * nx * ny * nz] - 1.0).reshape((nx, ny, nz), order="F")
    w = np.zeros((nx, ny, nz + 1), order="F")
    w[:, :, 1:nz] = (2.0 * d[2 * nx * ny * nz:2 * nx * ny * nz + nx * ny * (nz - 1)] - 1.0).reshape(
        (nx, ny, nz - 1), order="F")
    div = ((u - np.roll(u, 1, axis=0)) / h[0] + (v - np.roll(v, 1, axis=1)) / h[1]
           + (w[:, :, 1:] - w[:, :, :-1]) / h[2])
    return np.asfortranarray(div - div.mean())


def ensemble_sizes(count: int = 64, seed: int = SEED):
    """Subdomain sizes of the batched ensemble: nx, ny uniform in [32, 128], nz uniform in [32, 96]."""
    d = uniform(seed ^ 0xE45E, 3 * count)
    nx = 32 + (d[0::3] * 97).astype(int)
    ny = 32 + (d[1::3] * 97).astype(int)
    nz = 32 + (d[2::3] * 65).astype(int)
    return [(int(a), int(b), int(c)) for a, b, c in zip(nx, ny, nz)]
