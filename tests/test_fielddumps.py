"""Field dumps (fortran-to-cuda_amd/fielddumps.py): the storage-shape-bounds-v1 format that an
application's field-dump block writes (benchmarks/elmm/templates/field-dumps.f90), its reader checks
(containers/fields.py:7-34), and the pointwise and relative-L2 rules (containers/fields.py:37-66,
containers/README.md:161-175).

The format is pinned by a Fortran program that writes a dump exactly as the template does (stream
unformatted values, then a list-directed `storage_size, shape, lbound, ubound` line); the reader must
recover its values and bounds, and the writer must produce the same bytes and the same integers.
"""
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import __graft_entry__ as graft

fd = graft._module("fielddumps")

FORTRAN = """program dump
  implicit none
  real(8), allocatable :: U(:,:,:)
  integer :: u_unit, i, j, k
  allocate(U(-1:2, 0:3, 1:2))
  do k = 1, 2
    do j = 0, 3
      do i = -1, 2
        U(i, j, k) = 100*i + 10*j + k + 0.25d0
      end do
    end do
  end do
  open(newunit=u_unit, file='fields-U.bin', access='stream', form='unformatted', status='new')
  write(u_unit) U
  close(u_unit)
  open(newunit=u_unit, file='fields-U.shape', status='new')
  write(u_unit, *) storage_size(U), shape(U), lbound(U), ubound(U)
  close(u_unit)
end program
"""


def reference_array():
    i = np.arange(-1, 3)[:, None, None]
    j = np.arange(0, 4)[None, :, None]
    k = np.arange(1, 3)[None, None, :]
    return np.asfortranarray(100.0 * i + 10.0 * j + k + 0.25)


def fortran_compiler():
    return shutil.which("gfortran") or (Path("/opt/rocm/bin/amdflang") if Path("/opt/rocm/bin/amdflang").exists()
                                        else None)


def test_reader_and_writer_match_a_fortran_written_dump(tmp_path):
    fc = fortran_compiler()
    if fc is None:
        pytest.skip("no Fortran compiler")
    fdir = tmp_path / "fortran"
    fdir.mkdir()
    (fdir / "dump.f90").write_text(FORTRAN)
    subprocess.run([str(fc), "dump.f90", "-o", "dump"], cwd=fdir, check=True, capture_output=True)
    subprocess.run([str(fdir / "dump")], cwd=fdir, check=True)
    got = fd.read_dumps(fdir, np.float64, 1)
    values, layout = got["fields-U.bin"]
    want = reference_array()
    assert np.array_equal(values, want.ravel(order="F"))
    assert layout == {"bits": 64, "shape": [4, 4, 2], "lower_bounds": [-1, 0, 1], "upper_bounds": [2, 3, 2]}
    ours = tmp_path / "ours"
    fd.write_dump(ours, "U", want, lower_bounds=(-1, 0, 1))
    assert (ours / "fields-U.bin").read_bytes() == (fdir / "fields-U.bin").read_bytes()
    assert (ours / "fields-U.shape").read_text().split() == (fdir / "fields-U.shape").read_text().split()


def test_round_trip_of_several_ranks_and_precisions(tmp_path):
    rng = np.random.default_rng(3)
    arrays = {"a": rng.uniform(-1, 1, (5,)), "b": rng.uniform(-1, 1, (3, 4)).astype(np.float32),
              "c": np.zeros((2, 0, 3))}
    for name, a in arrays.items():
        d = tmp_path / name
        fd.write_dump(d, name, a, lower_bounds=[0] * a.ndim)
        (values, layout), = fd.read_dumps(d, a.dtype, 1).values()
        assert np.array_equal(values, np.asarray(a).ravel(order="F"))
        assert layout["shape"] == list(a.shape) and layout["lower_bounds"] == [0] * a.ndim
        assert layout["upper_bounds"] == [n - 1 for n in a.shape]


def test_reader_rejections(tmp_path):
    a = np.arange(12.0).reshape(3, 4)
    fd.write_dump(tmp_path, "A", a)
    with pytest.raises(fd.DumpError, match="expected 2 fields"):
        fd.read_dumps(tmp_path, np.float64, 2)
    with pytest.raises(fd.DumpError, match="precision"):
        fd.read_dumps(tmp_path, np.float32, 1)  # 32-bit values against a 64-bit storage size
    with pytest.raises(fd.DumpError, match="unsupported"):
        fd.read_dumps(tmp_path, np.float64, 1, metadata_format="v2")
    shape = tmp_path / "fields-A.shape"
    shape.write_text("64 3 4 1 1 3 5\n")  # upper bound 5 does not give extent 4
    with pytest.raises(fd.DumpError, match="extent"):
        fd.read_dumps(tmp_path, np.float64, 1)
    shape.write_text("64 3 5 1 1 3 5\n")  # 15 values claimed, 12 stored
    with pytest.raises(fd.DumpError, match="value count"):
        fd.read_dumps(tmp_path, np.float64, 1)
    shape.write_text("64 3 4 1 1\n")  # 3r + 1 integers required
    with pytest.raises(fd.DumpError, match="rank"):
        fd.read_dumps(tmp_path, np.float64, 1)
    (tmp_path / "fields-A.bin").write_bytes(b"\0" * 13)
    with pytest.raises(fd.DumpError, match="partial element"):
        fd.read_dumps(tmp_path, np.float64, 1, metadata_format=None)


def dumps(tmp_path, tag, **fields):
    d = tmp_path / tag
    for name, a in fields.items():
        fd.write_dump(d, name, a)
    return fd.read_dumps(d, np.float64, len(fields))


def test_pointwise_rule(tmp_path):
    r = np.linspace(-2.0, 2.0, 64).reshape(8, 8)
    ref = dumps(tmp_path, "ref", X=r, Y=r)
    a = r.copy()
    a[1, 2] += 1e-6 + 1e-5 * abs(r[1, 2]) * 0.5          # inside atol + rtol*|r|
    b = r.copy()
    b[3, 4] += 2e-6 + 1e-5 * abs(r[3, 4])                # outside
    got = fd.compare_pointwise(dumps(tmp_path, "cand", X=a, Y=b), ref)
    assert got["fields-X.bin"]["passed"] and got["fields-X.bin"]["failed_values"] == 0
    y = got["fields-Y.bin"]
    assert not y["passed"] and y["failed_values"] == 1 and y["failed_flat_indices"] == [3 + 8 * 4]
    assert y["layout"]["shape"] == [8, 8]
    # Nonfinite values fail, matching infinities and NaNs included, and the statistics stay finite JSON.
    inf = r.copy()
    inf[0, 0] = np.inf
    got = fd.compare_pointwise(dumps(tmp_path, "inf", X=inf, Y=r), dumps(tmp_path, "inf_ref", X=inf, Y=r))
    x = got["fields-X.bin"]
    assert not x["passed"] and x["candidate_nonfinite_values"] == 1 and x["reference_nonfinite_values"] == 1
    assert x["max_absolute_error"] is None
    with pytest.raises(fd.DumpError):
        fd.compare_pointwise(ref, ref, atol=-1.0)
    with pytest.raises(fd.DumpError):
        fd.compare_pointwise(dumps(tmp_path, "other", X=r), ref)


def test_relative_l2_rule(tmp_path):
    rng = np.random.default_rng(9)
    r = rng.uniform(-1, 1, (16, 16))
    noise = rng.uniform(-1, 1, r.shape)
    small = r + noise * (0.009 * np.linalg.norm(r) / np.linalg.norm(noise))
    large = r + noise * (0.011 * np.linalg.norm(r) / np.linalg.norm(noise))
    z = np.zeros((4, 4))
    zn = z.copy()
    zn[1, 1] = 2e-6
    got = fd.compare_l2(dumps(tmp_path, "c", A=small, B=large, Z=z + 5e-7, W=zn),
                        dumps(tmp_path, "r", A=r, B=r, Z=z, W=z))
    assert got["fields-A.bin"]["passed"] and math.isclose(got["fields-A.bin"]["relative_l2"], 0.009)
    assert not got["fields-B.bin"]["passed"]
    assert got["fields-Z.bin"]["passed"] and not got["fields-W.bin"]["passed"]  # zero reference: max |a| <= 1e-6


@pytest.mark.gpu
def test_les_step_dumps_of_the_device_match_the_oracle_pointwise(ftc, scope, hip, oracle, fields, tmp_path):
    # An application-level parity check through files: the resident step's fields (halos included), dumped
    # after observing them, against the oracle step's, with zero tolerances.
    nx, ny, nz, dx, dt, cs = 40, 24, 12, 10.0, 0.2, 0.17
    U, V, W, S = fields.les_fields(nx, ny, nz)
    src = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F"), S.copy(order="F")]
    dst = [np.zeros_like(U, order="F") for _ in range(4)]
    with scope.Domain(hip) as dom:
        hs = [dom.bind(x) for x in src]
        hd = [dom.bind(x, initialized=True) for x in dst]
        scope.check(hip, hip.cpp_fort_shared_les_step(dom.ctx, 1, *hd, *hs, ftc.dp(dx), ftc.dp(dx), ftc.dp(dx),
                                                       ftc.dp(dt), ftc.dp(cs), ftc.ip(nx), ftc.ip(ny), ftc.ip(nz),
                                                       ftc.ip(1)))
        dom.flush()
        for name, h, a in zip("UVWS", hd, dst):
            fd.dump_resident(dom, h, a, tmp_path / "device", name, lower_bounds=(0, 0, 0))
    for name, a in zip("UVWS", oracle.les_step(U, V, W, S, dx, dx, dx, dt, cs)):
        fd.write_dump(tmp_path / "oracle", name, a, lower_bounds=(0, 0, 0))
    got = fd.compare_pointwise(fd.read_dumps(tmp_path / "device", np.float64, 4),
                               fd.read_dumps(tmp_path / "oracle", np.float64, 4), atol=0.0, rtol=0.0)
    assert all(s["passed"] for s in got.values()), got
