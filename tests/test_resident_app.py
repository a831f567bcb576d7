"""Unchanged Fortran applications built from their TOML descriptions on this library's entries
(fortran-to-cuda_amd/resident_app.py), as the reference's resident flow builds them
(compiler/resident/application.py, directives.py, compiler/scopes/resident_splice.py).

On the CPU: the edits, the directive expansion and the procedure splice on a small application of this
file's own, compiled and run with the domain left inactive, so that every spliced block takes its
original statements; and the refusals. On the GPU: the reference's own CDU application
(benchmarks/cdu/cdu.toml), whose resident executable `make -C oracle apps` builds next to the native
one, must print the native checksum of U2 digit for digit at 64^3 and at a ragged 70 x 37 x 19; and the
reference's three stencil correctness drivers (compiler/tests/fixtures/stencils/*_main.f90, described by
oracle/apps/stencil_*.toml) must dump the interior of their output exactly as their native builds do, at
16^3 and at 23 x 9 x 5.
"""
import subprocess
from pathlib import Path

import pytest

import __graft_entry__ as graft

ROOT = Path(__file__).resolve().parents[1]
LAYER = ROOT / "oracle" / "_ref" / "resident"
APPS = ROOT / "oracle" / "_ref" / "apps"

app = graft._module("resident_app")

KERNEL = """module MomentumAdvection
  implicit none
  integer, parameter :: knd = kind(1d0)
contains
  subroutine helper(A, n)
    real(knd), intent(inout) :: A(:,:,:)
    integer, intent(in) :: n
    A(2:n+1, :, :) = 2 * A(2:n+1, :, :)
  end subroutine helper

  subroutine CDU(U2, U, V, W, dxmin, dymin, dzmin, Unx, Uny, Unz)
    real(knd), contiguous, intent(out) :: U2(:,:,:)
    real(knd), contiguous, intent(in) :: U(:,:,:), V(:,:,:), &
                                         W(:,:,:)
    real(knd), intent(in) :: dxmin, dymin, dzmin
    integer, intent(in) :: Unx, Uny, Unz
    ! a stand-in body: only which statements run matters here
    U2 = U + dxmin * V - dymin * W + dzmin
    call helper(U2, Unx)
  end subroutine CDU
end module MomentumAdvection
"""

MAIN = """program small
  use MomentumAdvection
  implicit none
  real(knd), allocatable :: U(:,:,:), V(:,:,:), W(:,:,:), U2(:,:,:)
  integer :: i
  allocate(U(5,4,3), V(5,4,3), W(5,4,3), U2(5,4,3))
  U = 1; V = 2; W = 3; U2 = 0
  ! ---- steps ----
  do i = 1, 2
    call CDU(U2, U, V, W, 0.5_knd, 0.25_knd, 1.0_knd, 3, 2, 1)
  end do
  print '(a,f0.4)', 'sum ', sum(U2)
end program small
"""

DESCRIPTION = """name = "small"
sources = ["src/kernel.f90", "src/main.f90"]
program = "src/main.f90"

[variants.default]
defines = []

[[selection]]
label = "advection"
procedures = ["momentumadvection::*"]

[[edit]]
file = "main.f90"
at = "  ! ---- steps ----\\n"
action = "before"
text = "  !$fort enter if(.false.) budget(1048576)\\n"

[[edit]]
file = "main.f90"
at = "  print '(a,f0.4)',"
action = "before"
text = "  !$fort observe(U2, U)\\n  !$fort leave\\n  !$fort publish if(i > 0)\\n"

[[edit]]
file = "main.f90"
at = "end    program small"
action = "replace"
text = "  !$fort close\\n{original}"
"""


def write_app(tmp_path, description=DESCRIPTION):
    (tmp_path / "src").mkdir(parents=True, exist_ok=True)
    (tmp_path / "src" / "kernel.f90").write_text(KERNEL)
    (tmp_path / "src" / "main.f90").write_text(MAIN)
    (tmp_path / "small.toml").write_text(description)
    return tmp_path / "small.toml"


def test_generated_sources_splice_the_entry_and_expand_every_directive(tmp_path):
    sources, chosen = app.generate(write_app(tmp_path), tmp_path / "gen")
    assert [p.name for p in sources] == ["kernel.f90", "main.f90"]
    assert chosen["spliced"] == [("kernel.f90", "momentumadvection::cdu")]
    assert chosen["native"] == [("kernel.f90", "momentumadvection::helper")]
    assert [r["directive"] for r in chosen["directives"]] == ["enter", "observe", "leave", "publish", "close"]
    kernel = sources[0].read_text()
    # The block follows the declarations (a continued one included) and wraps the original statements.
    head, block = kernel.split("subroutine CDU(", 1)[1].split("    block\n", 1)
    assert "W(:,:,:)" in head and "U2 = U" not in head
    assert "use fort_shared_momentumadvection_cdu, only: ftc_run => run" in block
    assert block.index("ftc_domain%publish()") < block.index("U2 = U + dxmin") < block.index("end block")
    for name in ("U2", "U", "V", "W"):
        assert f"ftc_domain%bind_real64({name}, lbound({name}), .true., &" in block
    assert "dxmin, dymin, dzmin, Unx, Uny, Unz)" in block
    main = sources[1].read_text()
    assert "!$fort" not in main
    assert "if (.false.) call ftc_domain%begin_epoch(int(0, c_int), int(1048576, c_size_t))" in main
    assert "if (ftc_domain%active .and. (i > 0)) call ftc_domain%publish()" in main
    assert main.rstrip().endswith("call ftc_domain%close()\n  end block\nend program small")


def test_an_inactive_domain_runs_the_original_statements(tmp_path):
    if not (LAYER / "resident_domain.o").exists():
        pytest.skip("the reference's Fortran host layer is not built (make -C oracle resident)")
    if not (ROOT / "fortran-to-cuda_amd" / "libftc_hip.so").exists():
        pytest.skip("libftc_hip.so is not built")
    exe = tmp_path / "bin" / "small"
    app.build(write_app(tmp_path), exe, LAYER)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    # U2 = 1 + 0.5*2 - 0.25*3 + 1 = 2.25 everywhere, doubled on rows 2..4: 60 values, 36 of them doubled.
    assert out.stdout.split() == ["sum", f"{2.25 * (60 + 36):.4f}"]


@pytest.mark.parametrize("change,message", [
    (("at = \"  ! ---- steps ----\\n\"", "at = \"  ! ---- nowhere ----\\n\""), "expected one"),
    (("action = \"before\"", "action = \"around\""), "before, after or replace"),
    (("!$fort enter if(.false.) budget(1048576)", "!$fort enter(U)"), "takes no argument list"),
    (("!$fort observe(U2, U)", "!$fort observe(U2(1,1,1))"), "not an array name"),
    (("!$fort leave", "!$fort leave now"), "unsupported clause 'now'"),
    (("!$fort close", "!$fort vanish"), "not a residency directive"),
])
def test_malformed_edits_and_directives_are_refused(tmp_path, change, message):
    with pytest.raises(ValueError, match=message):
        app.generate(write_app(tmp_path, DESCRIPTION.replace(*change)), tmp_path / "gen")


SOURCE = """module m
contains
  subroutine a(x)
    real :: x
    x = 1
    call b(x, &
           2)
  end subroutine a
  subroutine b(x, n)
    real :: x
    integer :: n
    x = n
    call c(x)
  end subroutine b
  real function f(x)
    real :: x
    f = x
  end function f
end module m
"""


def test_edit_locators(tmp_path):
    # compiler/resident/application.py:49-58, 262-290: procedure confinement, at, between, call (a continued
    # statement), module_end, the whole procedure, templates and {original}.
    edit = lambda **e: app.apply_edit(SOURCE, {"file": "m.f90", **e}, "edit", tmp_path)
    got = edit(procedure="b", at="x = n", action="replace", text="x = 2 * n")
    assert "x = 2 * n" in got and "x = 1" in got
    got = edit(call="b", action="before", text="    ! before\n")
    assert got.index("! before") < got.index("call b(x, &") and got.count("call b") == 1
    got = edit(call="b", action="after", text="    ! after\n")
    assert got.index("           2)\n    ! after") > 0
    got = edit(between=["x = 1", "2)"], action="replace", text="x = 3")
    assert "call b(x" not in got and "    x = 3\n  end subroutine a" in got
    got = edit(module_end="m", action="before", text="  ! last\n")
    assert got.rstrip().endswith("! last\nend module m")
    got = edit(procedure="f", function=True, action="replace", text="! gone\n")
    assert "function f" not in got and "! gone" in got
    (tmp_path / "t.f90").write_text("    ! {template}\n")
    got = edit(procedure="a", at="x = 1", action="replace", template="t.f90")
    assert "! {template}" in got
    got = edit(procedure="a", at="x = 1", action="replace", text="[{original}]")
    assert "[x = 1]" in got
    for bad, message in ((dict(call="c", procedure="a", action="after"), "expected one call of c"),
                         (dict(at="x", between=["a", "b"], action="after"), "one of at"),
                         (dict(procedure="z", action="after"), "expected one subroutine z"),
                         (dict(module_end="q", action="after"), "module q has no end module"),
                         (dict(action="after"), "needs at, between"),
                         (dict(at="x = 1", action="after", where="here"), "unknown edit keys")):
        with pytest.raises(ValueError, match=message):
            edit(**bad)


def test_an_entry_whose_procedure_has_other_arguments_is_refused(tmp_path):
    toml = write_app(tmp_path)
    kernel = tmp_path / "src" / "kernel.f90"
    kernel.write_text(KERNEL.replace("Unx, Uny, Unz)\n", "Unx, Uny)\n", 1))
    with pytest.raises(ValueError, match="has 9 arguments"):
        app.generate(toml, tmp_path / "gen")


def run_app(path):
    if not path.exists():
        pytest.skip(f"{path.name} is not built (make -C oracle apps)")
    out = subprocess.run([str(path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    return dict(line.split(":", 1) for line in out.stdout.splitlines() if ":" in line)


@pytest.mark.gpu
@pytest.mark.parametrize("resident,native", [("cdu", "../cdu_bench"), ("cdu_ragged", "cdu_native_ragged")])
def test_the_reference_cdu_application_prints_the_native_checksum(resident, native):
    got, want = run_app(APPS / resident), run_app(APPS / native)
    for key in ("grid", "iterations", "warmup_iters", "checksum"):
        assert got[key].split() == want[key].split(), key
    print(f"{resident}: ms_per_iter resident {got['ms_per_iter'].split()[0]}, native {want['ms_per_iter'].split()[0]}")


def run_output(path):
    if not path.exists():
        pytest.skip(f"{path.name} is not built (make -C oracle apps)")
    out = subprocess.run([str(path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout.split()


@pytest.mark.gpu
@pytest.mark.parametrize("stencil", ["cdu", "cvd", "cdw"])
@pytest.mark.parametrize("size", ["16", "ragged"])
def test_the_reference_stencil_drivers_dump_the_native_values(stencil, size):
    if size == "16":
        got, want = run_output(APPS / f"{stencil}_main"), run_output(ROOT / "oracle" / "_ref" / f"{stencil}_main")
        assert len(want) == 16 ** 3
    else:
        got, want = run_output(APPS / f"{stencil}_main_ragged"), run_output(APPS / f"{stencil}_main_native_ragged")
        assert len(want) == 23 * 9 * 5
    assert got == want  # g0.17, one value a line: the device values equal the native ones to the last digit
