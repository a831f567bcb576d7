"""The reference's relax application (compiler/tests/apps/relax/relax.f90: Jacobi sweeps written in one main
program, with array assignments and an external procedure) built resident on this library from
oracle/apps/relax.toml, against its native build (oracle/_ref/relax), as the reference's test compares them
(compiler/tests/test_resident_program_units_cuda.py:14-40): field.bin and every report line, byte for byte.

The builder splices a loop nest of a selected main program onto a hand-written entry when the hash of its
normalised text is listed (fortran-to-cuda_amd/resident_app.py, NEST_ENTRIES), and wraps the program's
other statements in the epoch that touch the nests' arrays in host-access blocks. On the CPU: the key
ignores layout, case, comments and OpenMP clauses; the parse, the accesses and the splice of a small
program of this file's own, compiled and run with the domain left inactive; and, where the reference tree
is present, that its three relax nests are the ones found. On the GPU: the run itself.
"""
import os
import re
import subprocess
from pathlib import Path

import pytest

import __graft_entry__ as graft

ROOT = Path(__file__).resolve().parents[1]
RESIDENT = ROOT / "oracle" / "_ref" / "apps" / "relax"
NATIVE = ROOT / "oracle" / "_ref" / "relax"
REFERENCE = Path("/root/reference/compiler/tests/apps/relax/relax.f90")

app = graft._module("resident_app")

PROGRAM = """program p
  implicit none
  real(8), allocatable :: a(:), b(:)
  real(8) :: s
  integer :: i, k
  allocate(a(8), b(8))
  a = 1
  !$fort enter
  do k = 1, 3
    !$omp parallel do
    do i = 1, 8
      b(i) = 2 * a(i)   ! doubled
    end do
    !$omp end parallel do
    s = 0
    if (k == 2) then
      a(1) = b(2) + 1
    else
      s = 1
    end if
    do i = 1, 8
      s = s + b(i)
    end do
  end do
  !$fort leave
  print *, s, a(1)
end program p
"""


def test_the_nest_key_ignores_layout_case_comments_and_openmp():
    one = ["!$omp parallel do private(i)\n", "do j = 1, n\n", "  do i = 1, n  ! inner\n",
           "    v(i,j) = 0.25 * (u(i-1,j) + &\n", "             u(i+1,j))\n", "  end do\n", "end do\n"]
    two = ["DO J=1,N\n", "DO I=1,N\n", "V(I,J)=0.25*(U(I-1,J)+U(I+1,J))\n", "END DO\n", "ENDDO\n"]
    assert app.nest_key(one) == app.nest_key(two)
    assert app.nest_key(one) != app.nest_key([l.replace("0.25", "0.5") for l in one])


def test_the_parse_and_the_accesses_of_a_program():
    lines = PROGRAM.splitlines(keepends=True)
    first, end = app.find_program(PROGRAM, "p")
    items = app._parse(lines, first, end)
    kinds = [(i.kind, lines[i.first].split()[0]) for i in items]
    assert kinds == [("statement", "allocate(a(8),"), ("statement", "a"), ("directive", "!$fort"),
                     ("construct", "do"), ("directive", "!$fort"), ("statement", "print")]
    loop = items[3]
    inner = loop.body[0]
    assert lines[inner.start].strip().startswith("!$omp parallel do")  # the directive belongs to the nest
    assert lines[inner.stop - 1].strip() == "!$omp end parallel do"
    branch = loop.body[2]
    assert [m.kind for m in branch.body] == ["statement", "middle", "statement"]
    assert app._accesses(lines, branch.body[0], {"a", "b"}) == (["b"], ["a"])
    assert app._accesses(lines, loop.body[3], {"a", "b"}) == (["b"], [])


def test_spliced_nests_and_guarded_statements_compile_and_run_on_the_host(tmp_path, monkeypatch):
    lines = PROGRAM.splitlines(keepends=True)
    first, end = app.find_program(PROGRAM, "p")
    items = app._parse(lines, first, end)
    key = app.nest_key(lines[items[3].body[0].first:items[3].body[0].stop])
    monkeypatch.setitem(app.NEST_ENTRIES, key, app.NestEntry("fort_shared_relax_sweep", ("b", "a"), ("k",)))
    text, records = app.splice_nests(PROGRAM, "p")
    assert [r["entry"] for r in records] == ["fort_shared_relax_sweep"]
    # The branch's statement that names a and b is guarded; the one that does not is left alone; the
    # reading loop is guarded whole.
    assert text.count("call ftc_domain%host_modify([ftc_domain%held(a)])") == 1
    assert text.count("call ftc_domain%observe([ftc_domain%held(b)])") == 2
    assert "  s = 1\n" in text and text.index("else") < text.index("  s = 1")
    layer = ROOT / "oracle" / "_ref" / "resident"
    if not (layer / "compiler").exists():
        pytest.skip("the reference's host layer is not built (make -C oracle resident)")
    src = tmp_path / "src"
    src.mkdir()
    (src / "p.f90").write_text(PROGRAM)
    toml = tmp_path / "p.toml"
    toml.write_text('name = "p"\nsources = ["p.f90"]\nprogram = "p.f90"\n[variants.default]\ndefines = []\n'
                    '[[selection]]\nlabel = "all"\nprocedures = ["$program::p"]\n')
    # Inactive domain: the enter directive is guarded off, so every block runs its original statements.
    inactive = PROGRAM.replace("!$fort enter", "!$fort enter if(.false.)")
    (src / "p.f90").write_text(inactive)
    exe = tmp_path / "p"
    app.build(toml, exe, layer, source_root=src)
    native = tmp_path / "native"
    fc = (layer / "compiler").read_text().strip()
    subprocess.run([fc, "-O2", "-o", str(native), str(src / "p.f90")], check=True, capture_output=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, check=True).stdout
    want = subprocess.run([str(native)], capture_output=True, text=True, timeout=60, check=True).stdout
    assert got.split() == want.split()


@pytest.mark.skipif(not REFERENCE.exists(), reason="the reference tree is not present")
def test_the_three_relax_nests_are_found_in_the_reference_program():
    text, records = app.splice_nests(REFERENCE.read_text(), "relax")
    assert [(r["line"], r["entry"]) for r in records] == [
        (30, "fort_shared_relax_sweep"), (38, "fort_shared_relax_change"), (44, "fort_shared_relax_copy")]
    # The host statement of every tenth sweep and the call of damp are guarded; nothing else is.
    assert text.count("ftc_host = ftc_domain%active") == 2


def run(binary, directory, *arguments, trace=False):
    if not binary.exists():
        pytest.skip(f"{binary.name} is not built (make -C oracle ref apps)")
    directory.mkdir()
    environment = {key: value for key, value in os.environ.items() if not key.startswith("FORT_")}
    environment.update(OMP_NUM_THREADS="4", OMP_DYNAMIC="FALSE")
    if trace:
        environment["FORT_RUNTIME_TRACE"] = "1"
    print(f"relax {binary.parent.name}/{binary.name} {' '.join(map(str, arguments))}", flush=True)
    result = subprocess.run([str(binary), *map(str, arguments)], cwd=directory, env=environment,
                            capture_output=True, text=True, timeout=120, check=False)
    assert result.returncode == 0, result.stdout[-2000:] + result.stderr[-4000:]
    return result.stdout, result.stderr, (directory / "field.bin").read_bytes()


@pytest.mark.gpu
@pytest.mark.parametrize("n,sweeps", [(96, 40), (33, 25)])
def test_the_resident_relax_run_matches_the_native_run_byte_for_byte(tmp_path, n, sweeps):
    native_output, _, native_field = run(NATIVE, tmp_path / "native", n, sweeps)
    output, trace, field = run(RESIDENT, tmp_path / "resident", n, sweeps, trace=True)
    assert len(field) == (n + 2) * (n + 2) * 8
    assert field == native_field
    # The largest change of a sweep comes back from the device digit for digit.
    assert output == native_output
    assert output.count("change") == sweeps // 10
    found = re.findall(r"^FORT_SCOPED close launches=(\d+)", trace, re.MULTILINE)
    # Per sweep: the stencil, the change and the copy on the device.
    assert found and int(found[-1]) >= 3 * sweeps, trace[-2000:]
