"""Complete-array field dumps: the format an application run writes for application-level parity, and
the two comparison rules applied to them.

Format (benchmarks/elmm/templates/field-dumps.f90; read back by containers/fields.py:7-34):
  fields-<name>.bin    the array's values in storage (column-major) order, native precision, no header;
  fields-<name>.shape  one list-directed line: storage_size in bits, then shape, lbound and ubound of the
                       rank-r array (3r + 1 integers), the "storage-shape-bounds-v1" metadata.

Rules:
  pointwise  per value |a - r| <= atol + rtol*|r| with both finite; a nonfinite value always fails,
             matching infinities included (containers/fields.py:37-66; atol 1e-6, rtol 1e-5 by default).
  l2         per field ||a - r||_2 <= 1% of ||r||_2; a field whose reference is all zero passes when
             max |a| <= 1e-6 (containers/README.md:161-175, the long-run rule).

A HIP run writes the same files from resident arrays after observing them (dump_resident), so a native
build of the reference and a run of this backend can be compared file for file.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

METADATA_FORMAT = "storage-shape-bounds-v1"
MAX_RANK = 15


class DumpError(ValueError):
    """A dump set that does not satisfy the format or the comparison preconditions."""


def write_dump(directory, name, array, lower_bounds=None):
    """Write fields-<name>.bin and fields-<name>.shape for `array` (any rank 1..15, numeric)."""
    a = np.asarray(array)
    if not 1 <= a.ndim <= MAX_RANK:
        raise DumpError(f"rank {a.ndim} outside 1..{MAX_RANK}")
    lower = [1] * a.ndim if lower_bounds is None else [int(v) for v in lower_bounds]
    if len(lower) != a.ndim:
        raise DumpError("one lower bound per dimension")
    upper = [lo + n - 1 for lo, n in zip(lower, a.shape)]
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    # Storage order of a Fortran array: column-major bytes, whatever the numpy array's own order.
    (d / f"fields-{name}.bin").write_bytes(np.asfortranarray(a).tobytes(order="F"))
    words = [a.dtype.itemsize * 8, *a.shape, *lower, *upper]
    (d / f"fields-{name}.shape").write_text(" ".join(f"{w:12d}" for w in words) + "\n")


def _layout(shape_path, itemsize, count):
    words = [int(w) for w in shape_path.read_text().split()]
    rank, extra = divmod(len(words) - 1, 3)
    if extra or not 1 <= rank <= MAX_RANK or words[0] != itemsize * 8:
        raise DumpError(f"{shape_path.name}: precision or metadata rank mismatch")
    shape, lower, upper = words[1:1 + rank], words[1 + rank:1 + 2 * rank], words[1 + 2 * rank:]
    for n, lo, hi in zip(shape, lower, upper):
        if n < 0 or n != max(0, hi - lo + 1):
            raise DumpError(f"{shape_path.name}: extent does not match its bounds")
    if math.prod(shape) != count:
        raise DumpError(f"{shape_path.name}: extents do not match the value count")
    return {"bits": words[0], "shape": shape, "lower_bounds": lower, "upper_bounds": upper}


def read_dumps(directory, dtype, field_count, metadata_format=METADATA_FORMAT):
    """All fields-*.bin of `directory` as {file name: (flat values, layout or None)}.

    metadata_format None reads the values only; METADATA_FORMAT also checks each .shape file."""
    if metadata_format not in (None, METADATA_FORMAT):
        raise DumpError(f"unsupported field metadata format {metadata_format!r}")
    dt = np.dtype(dtype)
    out = {}
    for path in sorted(Path(directory).glob("fields-*.bin")):
        size = path.stat().st_size
        if size % dt.itemsize:
            raise DumpError(f"{path.name}: partial element")
        values = np.fromfile(path, dtype=dt)
        layout = _layout(path.with_suffix(".shape"), dt.itemsize, values.size) if metadata_format else None
        out[path.name] = (values, layout)
    if len(out) != field_count:
        raise DumpError(f"expected {field_count} fields, found {len(out)}")
    return out


def _check_sets(actual, reference):
    if actual.keys() != reference.keys():
        raise DumpError("the two runs dumped different fields")
    for name, (values, layout) in actual.items():
        ref_values, ref_layout = reference[name]
        if values.shape != ref_values.shape or layout != ref_layout:
            raise DumpError(f"{name}: shape or original bounds differ")


def compare_pointwise(actual, reference, atol=1e-6, rtol=1e-5):
    """Per-field pointwise comparison of two read_dumps results; returns {name: statistics}."""
    for t in (atol, rtol):
        if not math.isfinite(t) or t < 0:
            raise DumpError("tolerances must be finite and nonnegative")
    _check_sets(actual, reference)
    result = {}
    for name, (a, layout) in actual.items():
        r = reference[name][0]
        fa, fr = np.isfinite(a), np.isfinite(r)
        with np.errstate(over="ignore", invalid="ignore"):
            err = np.abs(a - r)
            ok = fa & fr & (err <= atol + rtol * np.abs(r))
        bad = np.flatnonzero(~ok)
        ferr = np.isfinite(err)
        stats = {
            "values": int(a.size), "passed": bad.size == 0, "failed_values": int(bad.size),
            "candidate_nonfinite_values": int(a.size - np.count_nonzero(fa)),
            "reference_nonfinite_values": int(r.size - np.count_nonzero(fr)),
            "absolute_error_nonfinite_values": int(err.size - np.count_nonzero(ferr)),
            "max_absolute_error": float(err.max(initial=0.0)) if ferr.all() else None,
            "failed_flat_indices": bad[:8].tolist(), "failed_locations_truncated": bool(bad.size > 8),
            "atol": atol, "rtol": rtol,
        }
        if layout is not None:
            stats["layout"] = layout
        result[name] = stats
    return result


def compare_l2(actual, reference, relative=0.01, zero_atol=1e-6):
    """Per-field relative L2 comparison (the long-run rule); returns {name: statistics}."""
    _check_sets(actual, reference)
    result = {}
    for name, (a, _) in actual.items():
        r = reference[name][0]
        finite = bool(np.isfinite(a).all() and np.isfinite(r).all())
        ad, rd = a.astype(np.float64), r.astype(np.float64)
        rnorm = float(np.linalg.norm(rd)) if finite else math.nan
        if finite and rnorm == 0.0:
            worst = float(np.abs(ad).max(initial=0.0))
            result[name] = {"passed": worst <= zero_atol, "reference_norm": 0.0, "max_absolute_error": worst}
            continue
        rel = float(np.linalg.norm(ad - rd)) / rnorm if finite else math.nan
        result[name] = {"passed": finite and rel <= relative, "reference_norm": rnorm, "relative_l2": rel}
    return result


def dump_resident(domain, handle, array, directory, name, lower_bounds=None):
    """Observe a resident array (its host copy becomes current) and write it as a field dump."""
    domain.observe(handle)
    write_dump(directory, name, array, lower_bounds)
