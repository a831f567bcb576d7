"""The reference's box application (compiler/tests/apps/box: a periodic field diffused with point sources,
two halo layers, host and device diagnostics) built resident on this library's entries from
oracle/apps/box.toml, against its native build (oracle/_ref/box), as the reference's own tests compare
them (compiler/tests/test_resident_application_cuda.py:38-120): the field with its halos and what each
source saw (field.bin), and every report line, byte for byte.

The halo fill, the diffusion step and the store run on the device (fortran-to-cuda_amd/csrc/box.hip);
the sources and the summary and peak diagnostics stay on the host inside blocks that hand them the
current values. On the CPU: below the guarded extent the same executable never begins a device epoch and
runs its original loops. On the GPU: the report sizes of the reference's tests, a long run, sources that
share cells, an empty source list, and device budgets that hold both fields, one or neither.
"""
import os
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
RESIDENT = ROOT / "oracle" / "_ref" / "apps" / "box"
NATIVE = ROOT / "oracle" / "_ref" / "box"


def run(binary, directory, *arguments, trace=False, **variables):
    if not binary.exists():
        pytest.skip(f"{binary.name} is not built (make -C oracle ref apps)")
    directory.mkdir()
    environment = {key: value for key, value in os.environ.items() if not key.startswith("FORT_")}
    environment.update(OMP_NUM_THREADS="4", OMP_DYNAMIC="FALSE", **variables)
    if trace:
        environment["FORT_RUNTIME_TRACE"] = "1"
    print(f"box {binary.parent.name}/{binary.name} {' '.join(map(str, arguments))} {variables}", flush=True)
    result = subprocess.run([str(binary), *map(str, arguments)], cwd=directory, env=environment,
                            capture_output=True, text=True, timeout=120, check=False)
    assert result.returncode == 0, result.stdout[-2000:] + result.stderr[-4000:]
    return result.stdout, result.stderr, (directory / "field.bin").read_bytes()


def closing(trace):
    """The runtime's counters at close, from the traced run's last FORT_SCOPED close line."""
    found = re.findall(r"^FORT_SCOPED close (.*)$", trace, re.MULTILINE)
    return {k: int(v) for k, v in (item.split("=") for item in found[-1].split())} if found else None


def test_below_the_guarded_extent_the_same_binary_runs_its_original_loops(tmp_path):
    arguments = (1, 6, 6, 5, 1)
    output, trace, field = run(RESIDENT, tmp_path / "resident", *arguments, trace=True)
    native_output, _, native_field = run(NATIVE, tmp_path / "native", *arguments)
    assert (field, output) == (native_field, native_output)
    counters = closing(trace)
    assert counters is None or counters["launches"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("arguments", [(48, 40, 32, 30, 10), (64, 48, 40, 6, 3)])
def test_the_resident_run_matches_the_native_run_byte_for_byte(tmp_path, arguments):
    nx, ny, nz, steps, every = arguments
    native_output, _, native_field = run(NATIVE, tmp_path / "native", *arguments)
    output, trace, field = run(RESIDENT, tmp_path / "resident", *arguments, trace=True)
    assert field == native_field
    # The field with its two halo layers on every side, then what each of the five sources saw.
    assert len(field) == ((nx + 4) * (ny + 4) * (nz + 4) + 5) * 8
    # The host diagnostics saw the device's values at each report, digit for digit.
    assert output == native_output
    assert output.count("energy") == steps // every and output.count("peak") == steps // every
    # Each step runs the three ordered halo passes, the update and the store on the device.
    counters = closing(trace)
    assert counters is not None and counters["launches"] >= 5 * steps, trace[-2000:]


@pytest.mark.gpu
@pytest.mark.parametrize("count", [-6, 0, 6])
def test_shared_cells_and_empty_source_lists_give_the_native_result(tmp_path, count):
    arguments = (24, 20, 16, 12, 4, count)
    native_output, _, native_field = run(NATIVE, tmp_path / "native", *arguments)
    output, _, field = run(RESIDENT, tmp_path / "resident", *arguments)
    assert (field, output) == (native_field, native_output)
    assert len(field) == (28 * 24 * 20 + abs(count)) * 8


@pytest.mark.gpu
@pytest.mark.parametrize("budget,warned", [("1317988", False), ("700K", True), ("600K", True)])
def test_a_budget_that_is_too_small_costs_time_and_says_so_once(tmp_path, budget, warned):
    # 1317988 bytes hold the two fields (52 x 44 x 36 doubles each) and nothing else; 700K holds one, so
    # the diffusion and the store, which need both, run on the host; 600K holds neither.
    arguments = (48, 40, 32, 30, 10)
    output, _, field = run(RESIDENT, tmp_path / "roomy", *arguments)
    tight_output, trace, tight_field = run(RESIDENT, tmp_path / "tight", *arguments, trace=True,
                                           FORT_RESIDENT_DEVICE_BUDGET=budget)
    assert (tight_field, tight_output) == (field, output)
    notice = "Resident domain: the arrays in use exceed the device budget"
    assert trace.count(notice) == (1 if warned else 0)
    if warned:
        assert "loop nests whose arrays do not fit together run on the host" in trace
