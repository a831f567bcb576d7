import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import __graft_entry__ as graft  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device and libftc_hip.so")


@pytest.fixture(scope="session")
def ftc():
    return graft.package()


@pytest.fixture(scope="session")
def scope(ftc):
    return graft._module("scope")


@pytest.fixture(scope="session")
def fields(ftc):
    return graft._module("fields")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding

    binding.lib()
    return binding


@pytest.fixture(scope="session")
def hip(ftc):
    """The device library; GPU tests fail loudly (they do not skip) when it or the device is missing."""
    lib = ftc.load()
    assert lib.ftc_device_count() >= 1, "no HIP device visible to a GPU test"
    return lib


@pytest.fixture(scope="session")
def cpu_runtime(ftc):
    return ftc.load_cpu_backend()


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                                                 np.ascontiguousarray(b).view(np.uint64))


def describe_mismatch(got: np.ndarray, want: np.ndarray) -> str:
    bad = np.argwhere(got.view(np.uint64) != want.view(np.uint64))
    if not len(bad):
        return "identical"
    first = tuple(int(x) for x in bad[0])
    return f"{len(bad)} cells differ; first at {first}: got {got[first]!r} want {want[first]!r}"
