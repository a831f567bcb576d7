"""Pins the Poisson oracle (oracle/poisson.py) before the device solve is checked against it:

- manufactured solutions: products of the discrete eigenvectors (periodic cos/sin, Neumann cos) are
  eigenvectors of the FD2 operator, so b = lambda * phi has the exact answer phi;
- an independent FD2 stencil residual of the oracle's solution on a seeded divergence right-hand side;
- a sparse direct solve (scipy.sparse.linalg.spsolve) of the same FD2 system with the mean fixed by an
  extra constraint row: the oracle must agree to rounding.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import poisson


def eigen_mode(nx, ny, nz, h, p, q, r, bc):
    i, j, k = np.arange(nx), np.arange(ny), np.arange(nz)
    if bc[0] == "periodic":
        fx = np.cos(2 * np.pi * p * i / nx)
        lx = (2 * np.cos(2 * np.pi * p / nx) - 2) / h[0] ** 2
    else:
        fx = np.cos(np.pi * p * (i + 0.5) / nx)
        lx = (2 * np.cos(np.pi * p / nx) - 2) / h[0] ** 2
    if bc[1] == "periodic":
        fy = np.sin(2 * np.pi * q * j / ny)
        ly = (2 * np.cos(2 * np.pi * q / ny) - 2) / h[1] ** 2
    else:
        fy = np.cos(np.pi * q * (j + 0.5) / ny)
        ly = (2 * np.cos(np.pi * q / ny) - 2) / h[1] ** 2
    fz = np.cos(np.pi * r * (k + 0.5) / nz)
    lz = (2 * np.cos(np.pi * r / nz) - 2) / h[2] ** 2
    phi = fx[:, None, None] * fy[None, :, None] * fz[None, None, :]
    return phi, lx + ly + lz


@pytest.mark.parametrize("bc", [("periodic", "periodic"), ("neumann", "periodic"), ("periodic", "neumann")])
@pytest.mark.parametrize("mode", [(1, 2, 0), (3, 1, 2), (0, 4, 1), (2, 3, 4)])
def test_manufactured_eigenmode_is_solved_exactly(bc, mode):
    nx, ny, nz, h = 12, 10, 8, (10.0, 7.0, 3.0)
    phi, lam = eigen_mode(nx, ny, nz, h, *mode, bc)
    phi -= phi.mean()
    got = poisson.solve(lam * phi, h, bc)
    np.testing.assert_allclose(got, phi, atol=1e-12 * np.abs(phi).max())


def test_residual_of_a_divergence_rhs():
    nx, ny, nz, h = 24, 20, 16, (10.0, 10.0, 10.0)
    b = poisson.divergence_rhs(nx, ny, nz, h, seed=7)
    phi = poisson.solve(b, h)
    r = b - poisson.laplacian(phi, h)
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(b)
    assert abs(phi.mean()) <= 1e-14 * np.abs(phi).max()


def fd2_matrix(nx, ny, nz, h, bc):
    def axis(n, hd, kind):
        main = -2.0 * np.ones(n)
        off = np.ones(n - 1)
        m = sp.diags([off, main, off], [-1, 0, 1], format="lil")
        if kind == "periodic":
            m[0, n - 1] += 1.0
            m[n - 1, 0] += 1.0
        else:
            m[0, 0] += 1.0
            m[n - 1, n - 1] += 1.0
        return sp.csr_matrix(m) / (hd * hd)

    Ax, Ay, Az = axis(nx, h[0], bc[0]), axis(ny, h[1], bc[1]), axis(nz, h[2], "neumann")
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    # Fortran (i fastest) ordering: kron(z, kron(y, x)).
    return (sp.kron(Iz, sp.kron(Iy, Ax)) + sp.kron(Iz, sp.kron(Ay, Ix)) + sp.kron(Az, sp.kron(Iy, Ix))).tocsr()


@pytest.mark.parametrize("h", [(1.0, 2.0, 0.5), (1.0, 1.0, 1.0)])
@pytest.mark.parametrize("bc", [("periodic", "periodic"), ("neumann", "neumann")])
def test_oracle_agrees_with_a_sparse_direct_solve(bc, h):
    # Unit spacing in z makes the singular column's unshifted end pivot -r + r = 0: the solve must still
    # divide by no zero (errstate raises on any division by zero or invalid value).
    nx, ny, nz = 9, 7, 6
    b = poisson.divergence_rhs(nx, ny, nz, h, seed=3)
    A = fd2_matrix(nx, ny, nz, h, bc)
    n = nx * ny * nz
    # Border the singular system with the zero-mean constraint: [A 1; 1^T 0] [phi; mu] = [b; 0].
    ones = np.ones((n, 1))
    K = sp.bmat([[A, sp.csr_matrix(ones)], [sp.csr_matrix(ones.T), None]]).tocsc()
    sol = spla.spsolve(K, np.concatenate([b.ravel(order="F"), [0.0]]))
    want = sol[:n].reshape((nx, ny, nz), order="F")
    with np.errstate(all="raise"):
        got = poisson.solve(b, h, bc)
    np.testing.assert_allclose(got, want, atol=1e-11 * np.abs(want).max())
