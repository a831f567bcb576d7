"""The device pressure solve through the ELMM adapter and a fort_scope external transaction, against the
oracle (oracle/poisson.py, pinned in test_poisson_oracle.py).

Tolerances: the FP64 solve is direct (dense FP64 transforms + FP64 Thomas), so its field error is a few
ulps of the transforms: 1e-11 of max|phi|. The mixed solve refines with FP64 residuals until the relative
residual is <= 1e-12 and must then meet the field target of 1e-10 of max|phi|.
"""
import numpy as np
import pytest

from oracle import poisson as poisson_oracle

pytestmark = pytest.mark.gpu


def run(ftc, scope, hip, b, h, mixed, bc_xy=(0, 0), tol=1e-12, max_solves=8, measure=True):
    from __graft_entry__ import _module

    pressure = _module("pressure")
    nx, ny, nz = b.shape
    phi = np.zeros((nx + 4, ny + 4, nz + 4), order="F")
    phi[0, 0, 0] = 123.0  # a halo value the solve must preserve
    rhs = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    rhs[1:-1, 1:-1, 1:-1] = b
    with scope.Domain(hip) as dom:
        hp = dom.bind(phi, lower=(-1, -1, -1))
        hr = dom.bind(rhs, lower=(0, 0, 0))
        solver = pressure.PressureSolver(hip, dom, (nx, ny, nz), (nx * h[0], ny * h[1], nz * h[2]), mixed, bc_xy)
        solver.refinement(tol, max_solves, measure)
        ext = solver.solve(hp, hr)
        assert ext.started == 1 and ext.committed == 1
        report = solver.report()
        dom.observe(hp)
        solver.destroy()
    assert phi[0, 0, 0] == 123.0
    return phi[2:-2, 2:-2, 2:-2].copy(), report


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("n", [(64, 48, 32), (37, 29, 17), (128, 128, 64)])
def test_solve_matches_the_oracle(ftc, scope, hip, mixed, n):
    h = (10.0, 10.0, 10.0)
    b = poisson_oracle.divergence_rhs(*n, h, seed=sum(n))
    got, report = run(ftc, scope, hip, b, h, mixed)
    want = poisson_oracle.solve(b, h)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err <= (1e-10 if mixed else 1e-11), (err, report.iterations, report.relative_residual)
    assert report.relative_residual <= (1e-12 if mixed else 1e-12)
    if mixed:
        assert report.converged == 1 and report.iterations >= 2


@pytest.mark.parametrize("n,solves", [((64, 48, 32), 3), ((37, 29, 17), 3), ((64, 64, 32), 2), ((256, 128, 16), 2)])
def test_mixed_solve_with_a_fixed_solve_count(ftc, scope, hip, n, solves):
    # Applications without residual norms (no host synchronisation), as a captured step runs them: three
    # where the transforms are split-bf16 products (other sizes), two where they are the Hartley passes
    # with FP32 workspaces (periodic power-of-two x and y).
    h = (10.0, 10.0, 10.0)
    b = poisson_oracle.divergence_rhs(*n, h, seed=sum(n) + 1)
    got, report = run(ftc, scope, hip, b, h, True, max_solves=solves, measure=False)
    want = poisson_oracle.solve(b, h)
    assert report.iterations == solves
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("bc_xy", [(4, 0), (0, 4), (4, 4)])
def test_neumann_walls_in_x_or_y(ftc, scope, hip, bc_xy):
    n, h = (40, 36, 12), (2.0, 3.0, 1.0)
    kinds = tuple("periodic" if c == 0 else "neumann" for c in bc_xy)
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, n)
    b -= b.mean()
    got, _ = run(ftc, scope, hip, b, h, False, bc_xy)
    want = poisson_oracle.solve(b, h, kinds)
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


def test_manufactured_eigenmode(ftc, scope, hip):
    nx, ny, nz, h = 48, 32, 24, (1.0, 1.0, 1.0)
    i, j, k = np.arange(nx), np.arange(ny), np.arange(nz)
    phi = (np.cos(2 * np.pi * 3 * i / nx)[:, None, None] * np.sin(2 * np.pi * 2 * j / ny)[None, :, None]
           * np.cos(np.pi * 2 * (k + 0.5) / nz)[None, None, :])
    lam = ((2 * np.cos(2 * np.pi * 3 / nx) - 2) + (2 * np.cos(2 * np.pi * 2 / ny) - 2) + (2 * np.cos(np.pi * 2 / nz) - 2))
    got, _ = run(ftc, scope, hip, lam * phi, h, False)
    assert np.abs(got - phi).max() <= 1e-12


def test_unsupported_configurations_are_rejected(ftc, scope, hip):
    from __graft_entry__ import _module

    pressure = _module("pressure")
    with scope.Domain(hip) as dom:
        with pytest.raises(scope.ScopeError) as e:
            pressure.PressureSolver(hip, dom, (16, 16, 1), (1.0, 1.0, 1.0))
        assert e.value.status == scope.ARGUMENT


@pytest.mark.parametrize("n", [(256, 64, 16), (8, 16, 5), (4, 4, 3), (512, 512, 8), (4096, 8, 3)])
def test_fft_solve_matches_the_oracle_and_the_dense_solve(ftc, scope, hip, monkeypatch, n):
    # Periodic x and y of power-of-two sizes take the FFT (Hartley) transforms in the FP64 solve; with
    # FTC_POISSON_DENSE set at creation the same solver takes the dense FP64 MFMA transforms. Both are
    # direct solves of the same system, so they agree with the oracle and with each other to rounding.
    print(f"FFT solve {n}", flush=True)
    h = (10.0, 7.0, 5.0)
    b = poisson_oracle.divergence_rhs(*n, h, seed=sum(n) + 7)
    want = poisson_oracle.solve(b, h)
    fft, report = run(ftc, scope, hip, b, h, False)
    assert report.relative_residual <= 1e-12
    # The 4096-point workgroups (the y passes, and x passes of long lines) run 512 threads by default;
    # with 256 they hold twice the points per thread. Both are the same transform, operation for operation.
    monkeypatch.setenv("FTC_FFT_THREADS", "256")
    fft256, _ = run(ftc, scope, hip, b, h, False)
    monkeypatch.delenv("FTC_FFT_THREADS")
    assert np.array_equal(fft256, fft)
    monkeypatch.setenv("FTC_POISSON_DENSE", "1")
    dense, _ = run(ftc, scope, hip, b, h, False)
    scale = np.abs(want).max()
    assert np.abs(fft - want).max() <= 1e-11 * scale
    assert np.abs(fft - dense).max() <= 1e-11 * scale


@pytest.mark.parametrize("n", [(64, 32, 64), (32, 64, 96), (128, 64, 128), (64, 64, 48)])
def test_resident_z_solve_equals_the_chunked_sweep(ftc, scope, hip, monkeypatch, n):
    # The FP64 Hartley solve takes the z solve that holds whole columns in registers for 64, 96 or 128 levels
    # (48 levels keep the chunked sweep on both sides); it must give the chunked sweep's bits.
    h = (10.0, 10.0, 10.0)
    b = poisson_oracle.divergence_rhs(*n, h, seed=sum(n) + 7)
    resident, _ = run(ftc, scope, hip, b, h, False, max_solves=1, measure=False)
    monkeypatch.setenv("FTC_TRIDIAG_CHUNKED", "1")
    chunked, _ = run(ftc, scope, hip, b, h, False, max_solves=1, measure=False)
    assert np.array_equal(resident.view(np.uint64), chunked.view(np.uint64)), np.abs(resident - chunked).max()
    want = poisson_oracle.solve(b, h)
    assert np.abs(resident - want).max() / np.abs(want).max() <= 1e-11
