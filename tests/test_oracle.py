"""Pins the oracle's C restatement before anything is checked against it.

- CDU/CDV/CDW: bit equality with the reference's own Fortran (built from the reference sources into
  oracle/_ref by `make -C oracle ref`) on seeded inputs, and with the golden dumps of the unchanged
  reference drivers (tests/golden/stencil_16.npz);
- closed forms that a dropped term, a wrong sign or a swapped index would break: a uniform wind has zero
  tendencies; a wind linear in one axis gives a hand-computed CDU; the scalar flux form conserves the sum;
- Smagorinsky: zero for a uniform wind, (Cs*Delta)^2*gamma for the shear u = gamma*z, nonnegative;
  constant viscosity reduces the momentum diffusion to the box 7-point Laplacian
  (compiler/tests/apps/box/stepper.f90:8-15) to rounding;
- halos: every ghost equals its image (brute force).
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import bits_equal, describe_mismatch

GOLDEN = Path(__file__).resolve().parent / "golden" / "stencil_16.npz"


def random_fields(shape, seed):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.uniform(-1.0, 1.0, shape)) for _ in range(4)]


@pytest.mark.parametrize("name", ["cdu", "cdv", "cdw"])
def test_oracle_matches_the_golden_dump_of_the_reference_driver(oracle, name):
    g = np.load(GOLDEN)
    U, V, W = g["U"], g["V"], g["W"]
    n = 16
    out = oracle.component(name, np.asfortranarray(U), np.asfortranarray(V), np.asfortranarray(W),
                           1.0 / n, 1.0 / n, 1.0 / n, n, n, n)
    want = g[{"cdu": "U2", "cdv": "V2", "cdw": "W2"}[name]]
    got = np.ascontiguousarray(out[1:-1, 1:-1, 1:-1])
    assert bits_equal(got, np.ascontiguousarray(want)), describe_mismatch(got, want)


@pytest.mark.parametrize("name", ["cdu", "cdv", "cdw"])
@pytest.mark.parametrize("shape,loops", [((18, 18, 18), (16, 16, 16)), ((23, 9, 14), (21, 7, 12)),
                                         ((30, 20, 10), (25, 13, 8))])
def test_oracle_matches_the_reference_fortran(oracle, name, shape, loops):
    ref = oracle.reference_component(name)
    if ref is None:
        pytest.skip("oracle/_ref was not built (needs the reference tree and a Fortran compiler)")
    U, V, W, _ = random_fields(shape, 7 + sum(shape))
    dx, dy, dz = 0.37, 1.9, 0.61
    got = oracle.component(name, U, V, W, dx, dy, dz, *loops)
    want = ref(U, V, W, dx, dy, dz, *loops)
    assert bits_equal(got, want), describe_mismatch(got, want)


def test_uniform_wind_has_zero_tendencies(oracle):
    shape = (14, 12, 10)
    U = np.full(shape, 5.0, order="F")
    V = np.full(shape, -2.0, order="F")
    W = np.zeros(shape, order="F")
    S = np.full(shape, 0.3, order="F")
    for cs in (None, 0.17):
        for t in oracle.les_tendency(U, V, W, S, 10.0, 10.0, 10.0, cs):
            assert np.all(t == 0.0)


def test_cdu_of_a_wind_linear_in_x_is_the_hand_computed_value(oracle):
    # U = a*x (x = i), V = W = 0: CDUdiv gives -Ax*((2u+a)^2 - (2u-a)^2) = -8*Ax*a*u,
    # CDUadv gives -(Bx*2a*u); with Ax = 0.25/dx, Bx = 0.5/dx the tendency is 0.5*(-2au/dx - au/dx).
    shape = (12, 8, 8)
    a, dx = 0.5, 2.0
    i = np.arange(1, shape[0] + 1, dtype=np.float64)
    U = np.asfortranarray(np.broadcast_to((a * i)[:, None, None], shape).copy())
    Z = np.zeros(shape, order="F")
    T = oracle.component("cdu", U, Z, Z, dx, 1.0, 1.0, 10, 6, 6)
    u = U[1:-1, 1:-1, 1:-1]
    expect = 0.5 * (-(2.0 * a * u / dx) - (a * u / dx))
    np.testing.assert_allclose(T[1:-1, 1:-1, 1:-1], expect, rtol=1e-14, atol=0)


def test_scalar_advection_conserves_the_scalar(oracle, fields):
    U, V, W, S = fields.les_fields(24, 20, 12)
    _, _, _, TS = oracle.les_tendency(U, V, W, S, 10.0, 10.0, 10.0, None)
    total = TS[1:-1, 1:-1, 1:-1].sum()
    scale = np.abs(TS[1:-1, 1:-1, 1:-1]).sum()
    assert abs(total) <= 1e-12 * scale


def test_scalar_diffusion_conserves_the_scalar(oracle, fields):
    U, V, W, S = fields.les_fields(20, 16, 10)
    adv = oracle.les_tendency(U, V, W, S, 10.0, 10.0, 10.0, None)[3]
    full = oracle.les_tendency(U, V, W, S, 10.0, 10.0, 10.0, 0.17)[3]
    diff = (full - adv)[1:-1, 1:-1, 1:-1]
    assert abs(diff.sum()) <= 1e-12 * np.abs(diff).sum()


def test_smagorinsky_of_a_linear_shear(oracle):
    # u = gamma * z: |S| = sqrt(2*2*(gamma/2)^2) = gamma, so nu = (Cs*Delta)^2 * gamma away from the walls.
    nx, ny, nz, d, gamma, cs = 8, 6, 10, 10.0, 0.01, 0.17
    shape = (nx + 2, ny + 2, nz + 2)
    zc = (np.arange(nz + 2) - 0.5) * d
    U = np.asfortranarray(np.broadcast_to((gamma * zc)[None, None, :], shape).copy())
    V = np.zeros(shape, order="F")
    W = np.zeros(shape, order="F")
    NU = oracle.sgs_nu(U, V, W, d, d, d, cs)
    delta = np.cbrt(d * d * d)
    inner = NU[1:-1, 1:-1, 2:-2]
    np.testing.assert_allclose(inner, (cs * delta) ** 2 * gamma, rtol=1e-12)
    assert np.all(NU >= 0)
    assert np.all(oracle.sgs_nu(np.full(shape, 3.0, order="F"), V, W, d, d, d, cs) == 0.0)


def test_constant_viscosity_diffusion_is_the_box_laplacian(oracle):
    # With nu forced to a constant (a linear shear in w gives a uniform |S|), DU = nu * 7-point Laplacian.
    rng = np.random.default_rng(3)
    nx, ny, nz, d, cs = 10, 8, 6, 1.0, 0.2
    shape = (nx + 2, ny + 2, nz + 2)
    U = np.asfortranarray(rng.uniform(-1e-13, 1e-13, shape))
    x = (np.arange(nx + 2) - 0.5) * d
    # w = gamma*x has dw/dx = gamma everywhere except across the periodic seam; restrict checks inside.
    W = np.asfortranarray(np.broadcast_to((0.1 * x)[:, None, None], shape).copy())
    V = np.zeros(shape, order="F")
    S = np.zeros(shape, order="F")
    NU = oracle.sgs_nu(U, V, W, d, d, d, cs)
    nu0 = NU[4, 4, 3]
    adv = oracle.les_tendency(U, V, W, S, d, d, d, None)[0]
    full = oracle.les_tendency(U, V, W, S, d, d, d, cs)[0]
    DU = full - adv
    sl = (slice(3, nx - 1), slice(2, ny), slice(2, nz))
    c = U[sl]
    lap = (U[3 - 1:nx - 2, 2:ny, 2:nz] + U[3 + 1:nx, 2:ny, 2:nz] + U[3:nx - 1, 1:ny - 1, 2:nz]
           + U[3:nx - 1, 3:ny + 1, 2:nz] + U[3:nx - 1, 2:ny, 1:nz - 1] + U[3:nx - 1, 2:ny, 3:nz + 1]
           - 6 * c) / (d * d)
    assert np.allclose(NU[2:nx, 1:ny + 1, 1:nz + 1], nu0, rtol=1e-12)
    np.testing.assert_allclose(DU[sl], nu0 * lap, rtol=1e-9, atol=1e-28)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_halo_ghosts_equal_their_images(oracle, kind):
    nx, ny, nz = 7, 5, 4
    rng = np.random.default_rng(kind)
    F = np.asfortranarray(rng.uniform(-1, 1, (nx + 2, ny + 2, nz + 2)))
    oracle.halo_les(F, kind)
    for i in range(nx + 2):
        for j in range(ny + 2):
            for k in range(nz + 2):
                ii = nx if i == 0 else (1 if i == nx + 1 else i)
                jj = ny if j == 0 else (1 if j == ny + 1 else j)
                if kind == 1 and k in (0, nz, nz + 1):
                    assert F[i, j, k] == 0.0
                    continue
                kk, sign = k, 1.0
                if k == 0:
                    kk, sign = 1, (-1.0 if kind == 0 else 1.0)
                elif k == nz + 1:
                    kk, sign = nz, (-1.0 if kind == 0 else 1.0)
                assert F[i, j, k] == sign * F[ii, jj, kk]


def test_generated_fields_have_consistent_halos(oracle, fields):
    U, V, W, S = fields.les_fields(12, 10, 8)
    for F, kind in ((U, 0), (V, 0), (W, 1), (S, 2)):
        G = F.copy(order="F")
        oracle.halo_les(G, kind)
        assert bits_equal(F, G)


BOX = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "box"


def run_box(directory, nx, ny, nz, steps):
    # compiler/tests/apps/box/main.f90: box NX NY NZ STEPS REPORT_EVERY SOURCES writes u(-1:nx+2,...)
    # and then what each source saw to field.bin; no sources here.
    directory.mkdir()
    env = {**os.environ, "OMP_NUM_THREADS": "4", "OMP_DYNAMIC": "FALSE"}
    subprocess.run([str(BOX), str(nx), str(ny), str(nz), str(steps), "1000000", "0"], cwd=directory, env=env,
                   check=True, capture_output=True, timeout=120)
    raw = np.fromfile(directory / "field.bin", dtype=np.float64)
    assert raw.size == (nx + 4) * (ny + 4) * (nz + 4)
    return np.asfortranarray(raw.reshape((nx + 4, ny + 4, nz + 4), order="F"))


@pytest.mark.parametrize("nx,ny,nz,steps", [(12, 10, 8, 5), (7, 5, 3, 4), (48, 40, 32, 3)])
def test_box_oracle_reproduces_the_reference_box_field_bit_for_bit(oracle, tmp_path, nx, ny, nz, steps):
    # The reference's explicit diffusion with its two-layer periodic halo (compiler/tests/apps/box/
    # stepper.f90:8-37, boundaries.f90:6-41), built natively from its own sources into oracle/_ref/box: the
    # field after `steps` steps equals the oracle's steps applied to the field the same program writes
    # after none, every cell including the halos.
    if not BOX.exists():
        pytest.skip("oracle/_ref/box is not built (make -C oracle ref)")
    start = run_box(tmp_path / "start", nx, ny, nz, 0)
    want = run_box(tmp_path / "end", nx, ny, nz, steps)
    dx = 1.0 / nx
    nu = 0.1
    dt = 0.1 * dx * dx / nu
    a, b = start.copy(order="F"), np.zeros_like(start, order="F")
    for _ in range(steps):
        oracle.box_step(a, b, dx, dt, nu)
    assert bits_equal(a, want), describe_mismatch(a, want)
    # The halo after the last step holds the images taken at its start, so it differs from the interior's
    # images now: the comparison above covers the halo as the program leaves it.
    assert not bits_equal(a[2:-2, 2:-2, 2:-2], start[2:-2, 2:-2, 2:-2])
