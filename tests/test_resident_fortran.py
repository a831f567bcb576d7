"""Programs written against the reference's own Fortran host layer drive this library.

The module fort_resident_domain (compiler/runtime/resident_domain.f90) and the bind(C) interfaces and
structures of fort_scoped_memory (compiler/runtime/scoped_memory.f90) are compiled as they are, by
`make -C oracle resident`, into oracle/_ref/resident/; a program that uses them links against
libftc_cpu.so or libftc_hip.so in place of the reference's scoped_runtime.cu, the way the reference's
own tests of that layer link it (compiler/tests/test_resident_host_threads_cuda.py:53-72). Every symbol
the layer names must resolve, every structure it passes must have the layout it declares, and its
statuses and messages must come back as the reference's runtime gives them.

tests/fortran/resident_host.f90 runs on the CPU backend; tests/fortran/resident_entries.f90 calls the
LES step and chain contraction entries through the interface modules this library ships
(fortran-to-cuda_amd/fortran/ftc_entries.f90) as spliced application code calls them, and exchanges its
arrays as field dumps, which are compared with the oracle (bit for bit) and with numpy's FP64 products.
tests/fortran/resident_pressure.f90 runs ELMM's own resident pressure state
(benchmarks/elmm/sources/resident_pressure_state.f90 and gpu_pressure.f90, compiled against this library's
poisson_solver_device module in place of the absent poisson-solver submodule's) on this library's
pressure provider, against the oracle's direct solve.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import __graft_entry__ as graft

ROOT = Path(__file__).resolve().parents[1]
LAYER = ROOT / "oracle" / "_ref" / "resident"
SOURCES = Path(__file__).resolve().parent / "fortran"
PACKAGE = ROOT / "fortran-to-cuda_amd"

fd = graft._module("fielddumps")


def build(tmp_path, source, library, extra=()):
    """Compiles `source` against the layer's modules and links it to `library` (ftc_cpu or ftc_hip), with
    the layer objects `extra` besides the two host-layer modules."""
    compiler_file = LAYER / "compiler"
    if not (LAYER / "resident_domain.o").exists() or not compiler_file.exists():
        pytest.skip("the reference's Fortran host layer is not built (make -C oracle resident)")
    fc = compiler_file.read_text().strip()
    if not Path(fc).exists():
        pytest.skip(f"{fc} unavailable")
    if not (PACKAGE / f"lib{library}.so").exists():
        pytest.skip(f"lib{library}.so is not built")
    run = lambda cmd: subprocess.run(cmd, cwd=tmp_path, check=True, capture_output=True, text=True, timeout=120)
    objects = []
    # The entries' interface modules this library ships come first.
    for src in (PACKAGE / "fortran" / "ftc_entries.f90", source):
        obj = tmp_path / (src.stem + ".o")
        run([fc, "-cpp", "-O2", "-fopenmp", "-I", str(LAYER), "-module-dir", str(tmp_path), "-c", str(src),
             "-o", str(obj)])
        objects.append(str(obj))
    binary = tmp_path / source.stem
    for name in extra:
        if not (LAYER / name).exists():
            pytest.skip(f"{name} is not built (make -C oracle resident)")
    run([fc, "-fopenmp", *objects, *(str(LAYER / name) for name in extra), str(LAYER / "scoped_memory.o"),
         str(LAYER / "resident_domain.o"),
         "-L", str(PACKAGE), f"-l{library}", f"-Wl,-rpath,{PACKAGE}", "-o", str(binary)])
    return binary


def lines(text):
    return {line.split()[0]: line.split()[1:] for line in text.splitlines() if line.strip()}


@pytest.fixture(scope="module")
def host_program(tmp_path_factory):
    return build(tmp_path_factory.mktemp("resident_host"), SOURCES / "resident_host.f90", "ftc_cpu")


def run_host(program, case):
    return subprocess.run([str(program), case], capture_output=True, text=True, timeout=60)


def test_epochs_bindings_transfers_mirrors_and_workspace(host_program):
    result = run_host(host_program, "epochs")
    assert result.returncode == 0, result.stdout + result.stderr
    got = lines(result.stdout)
    assert result.stdout.split()[:2] == ["generation", "1"] and got["generation"] == ["2"]
    assert got["held"] == ["T", "T", "T"] and got["distinct"] == ["T"] and got["rebind"] == ["T"]
    # A device overwrite uploads nothing and reaches the host only when observed.
    assert got["host_before_observe"] == ["300.0"]
    assert got["host_after_observe"] == [f"{sum(100.0 + i for i in range(1, 25)):.1f}"]
    assert got["device_after_host_change"] == ["2724.0"]
    # Exactly one download (the observation) and one upload (the read after the host change), 24 doubles each.
    assert got["transfers"] == ["1", "1", "192", "192"]
    # A field of an array of derived type travels through its own copy; the other field is untouched.
    assert got["field_on_device"] == ["10.0", "20.0", "30.0"]
    assert got["field_before_observe"] == ["10.0", "20.0", "30.0"]
    assert got["field_after_observe"] == ["10.5", "20.5", "30.5", "1", "2", "3"]
    assert got["work_bounds"] == ["4", "3", "2"]
    # The epoch's end publishes and forgets its bindings; a workspace slot outlives the epoch.
    assert got["published"] == ["2724.0", "T"] and got["work_reused"] == ["T"]
    assert got["closed"] == []


def test_an_undefined_device_read_is_refused_with_its_message(host_program):
    result = run_host(host_program, "undefined")
    assert result.returncode == 0, result.stdout + result.stderr
    got = lines(result.stdout)
    assert got["status"] == ["7", "T"]  # FORT_SCOPE_UNINITIALIZED as the Fortran layer declares it
    assert "never defined" in " ".join(got["message"])


def test_a_foreign_handle_stops_the_program_through_the_layer(host_program):
    result = run_host(host_program, "foreign")
    assert result.returncode != 0
    assert "Resident runtime failed: STALE (status 2)" in result.stderr
    assert "Runtime: unknown or retired buffer handle" in result.stderr


@pytest.fixture(scope="module")
def entries_program(tmp_path_factory):
    return build(tmp_path_factory.mktemp("resident_entries"), SOURCES / "resident_entries.f90", "ftc_hip")


def run_entries(program, workdir, *args):
    (workdir / "out").mkdir(exist_ok=True)
    result = subprocess.run([str(program), *map(str, args)], cwd=workdir, capture_output=True, text=True,
                            timeout=120)
    assert result.returncode == 0, result.stdout + result.stderr
    assert result.stdout.split() == ["done"]


@pytest.mark.gpu
@pytest.mark.parametrize("n,steps", [((40, 24, 12), 3), ((17, 9, 5), 4)])
def test_les_steps_called_from_fortran_match_the_oracle_bit_for_bit(entries_program, oracle, fields, tmp_path,
                                                                    n, steps):
    nx, ny, nz = n
    dx, dt, cs = 10.0, 0.2, 0.17
    U, V, W, S = fields.les_fields(nx, ny, nz)
    for name, a in zip("UVWS", (U, V, W, S)):
        fd.write_dump(tmp_path / "in", name, a, lower_bounds=(0, 0, 0))
    run_entries(entries_program, tmp_path, "les", nx, ny, nz, steps, repr(dx), repr(dt), repr(cs))
    want = [U, V, W, S]
    for _ in range(steps):
        want = oracle.les_step(*want, dx, dx, dx, dt, cs)
    wdir = tmp_path / "want"
    for name, a in zip("UVWS", want):
        fd.write_dump(wdir, name, a, lower_bounds=(0, 0, 0))
    got = fd.compare_pointwise(fd.read_dumps(tmp_path / "out", np.float64, 4),
                               fd.read_dumps(wdir, np.float64, 4), atol=0.0, rtol=0.0)
    assert all(s["passed"] for s in got.values()), got
    for name in "UVWS":
        assert (tmp_path / "out" / f"fields-{name}.bin").read_bytes() == (wdir / f"fields-{name}.bin").read_bytes()


@pytest.mark.gpu
@pytest.mark.parametrize("n,precision,tol", [(96, 0, 1e-13), (128, 1, 2e-4)])
def test_chain_contraction_called_from_fortran(entries_program, tmp_path, n, precision, tol):
    rng = np.random.default_rng(41 + precision)
    A, B, C = (np.asfortranarray(rng.uniform(-1, 1, (n, n))) for _ in range(3))
    for name, a in zip("ABC", (A, B, C)):
        fd.write_dump(tmp_path / "in", name, a)
    run_entries(entries_program, tmp_path, "chain", n, precision)
    (values, layout), = fd.read_dumps(tmp_path / "out", np.float64, 1).values()
    assert layout == {"bits": 64, "shape": [n, n], "lower_bounds": [1, 1], "upper_bounds": [n, n]}
    D = values.reshape((n, n), order="F")
    want = A @ B @ C
    assert np.linalg.norm(D - want) <= tol * np.linalg.norm(want)


@pytest.fixture(scope="module")
def pressure_program(tmp_path_factory):
    return build(tmp_path_factory.mktemp("resident_pressure"), SOURCES / "resident_pressure.f90", "ftc_hip",
                 extra=("poisson_solver_device.o", "gpu_pressure.o", "resident_pressure_state.o"))


@pytest.mark.gpu
@pytest.mark.parametrize("flags,tol", [(0, 1e-11), (2, 1e-10)])
def test_elmm_resident_pressure_state_solves_on_this_library(pressure_program, tmp_path, flags, tol):
    from oracle import poisson as poisson_oracle

    n, h = (37, 29, 17), (10.0, 10.0, 10.0)
    b = poisson_oracle.divergence_rhs(*n, h, seed=11)
    rhs = np.zeros(tuple(m + 2 for m in n), order="F")
    rhs[1:-1, 1:-1, 1:-1] = b
    fd.write_dump(tmp_path / "in", "RHS", rhs, lower_bounds=(0, 0, 0))
    run_entries(pressure_program, tmp_path, *n, repr(h[0]), flags)
    (values, layout), = fd.read_dumps(tmp_path / "out", np.float64, 1).values()
    assert layout["lower_bounds"] == [-1, -1, -1] and layout["shape"] == [m + 4 for m in n]
    phi = values.reshape(layout["shape"], order="F")
    assert phi[0, 0, 0] == 123.0  # the ghost the driver planted is left alone
    want = poisson_oracle.solve(b, h)
    err = np.abs(phi[2:-2, 2:-2, 2:-2] - want).max() / np.abs(want).max()
    assert err <= tol, err
