"""The native Fortran build of the synthetic LES step (oracle/native/les_native.f90, built into
oracle/_build/les_native with -O3 -fopenmp -ffp-contract=off), which bench.py times as the CPU baseline,
gives the C restatement's step (oracle/step.c) bit for bit, halos included."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import bits_equal, describe_mismatch

NATIVE = Path(__file__).resolve().parents[1] / "oracle" / "_build" / "les_native"


@pytest.mark.parametrize("n", [(40, 24, 16), (23, 9, 5), (64, 64, 32), (2, 3, 2)])
def test_native_step_equals_the_c_restatement_bit_for_bit(oracle, fields, tmp_path, n):
    if not NATIVE.exists():
        pytest.skip("oracle/_build/les_native is not built (make -C oracle)")
    nx, ny, nz = n
    state = fields.les_fields(nx, ny, nz)
    np.concatenate([a.ravel(order="F") for a in state]).tofile(tmp_path / "in.bin")
    run = subprocess.run([str(NATIVE), str(nx), str(ny), str(nz), "1", "1", str(tmp_path / "in.bin"),
                          str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120,
                         env={"OMP_NUM_THREADS": "4", "PATH": "/usr/bin:/bin"})
    assert run.returncode == 0, run.stderr
    assert "median_ms_per_step" in run.stdout
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    size = (nx + 2) * (ny + 2) * (nz + 2)
    assert raw.size == 4 * size
    got = [raw[f * size:(f + 1) * size].reshape((nx + 2, ny + 2, nz + 2), order="F") for f in range(4)]
    want = oracle.les_step(*state, 10.0, 10.0, 10.0, 0.2, 0.17)
    for name, g, w in zip("UVWS", got, want):
        assert bits_equal(g, w), f"{name}: " + describe_mismatch(g, w)


def test_native_final_fields_after_several_steps_equal_the_c_restatement(oracle, fields, tmp_path):
    # les_native NX NY NZ STEPS REPEATS INPUT OUTPUT FINAL: FINAL holds the fields after 1 + REPEATS*STEPS steps.
    if not NATIVE.exists():
        pytest.skip("oracle/_build/les_native is not built (make -C oracle)")
    nx, ny, nz = 24, 20, 12
    state = fields.les_fields(nx, ny, nz)
    np.concatenate([a.ravel(order="F") for a in state]).tofile(tmp_path / "in.bin")
    run = subprocess.run([str(NATIVE), str(nx), str(ny), str(nz), "2", "2", str(tmp_path / "in.bin"),
                          str(tmp_path / "out.bin"), str(tmp_path / "final.bin")], capture_output=True, text=True,
                         timeout=120, env={"OMP_NUM_THREADS": "4", "PATH": "/usr/bin:/bin"})
    assert run.returncode == 0, run.stderr
    raw = np.fromfile(tmp_path / "final.bin", dtype=np.float64)
    size = (nx + 2) * (ny + 2) * (nz + 2)
    got = [raw[f * size:(f + 1) * size].reshape((nx + 2, ny + 2, nz + 2), order="F") for f in range(4)]
    want = list(state)
    for _ in range(5):
        want = oracle.les_step(*want, 10.0, 10.0, 10.0, 0.2, 0.17)
    for name, g, w in zip("UVWS", got, want):
        assert bits_equal(g, w), f"{name}: " + describe_mismatch(g, w)
