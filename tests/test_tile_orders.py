"""Every tile order the kernels use covers every tile exactly once (csrc/xcd.hpp: the launch order and the
XCD-contiguous order of the LES step, the persistent banded walk, the contraction order), counted on the
device by ftc_tile_visit_probe_v1, as the reference counts the visits of its grid-stride decode
(compiler/tests/test_cuda_coordinate_quotient.py). Shapes include single tiles, ragged counts that do not
divide by the eight XCDs, the bench's LES grid (8 x 128 x 3) and walks by fewer and by more workgroups than
there are tiles.
"""
import ctypes

import numpy as np
import pytest

GRIDS = [(1, 1, 1), (7, 5, 3), (13, 1, 1), (3, 17, 2), (8, 128, 3), (64, 64, 4), (9, 9, 9), (32, 32, 1), (17, 8, 2)]


def visits(hip, scheme, grid, workgroups=0):
    gx, gy, gz = grid
    counts = (ctypes.c_uint * (gx * gy * gz + 1))()
    assert hip.ftc_tile_visit_probe_v1(scheme, gx, gy, gz, workgroups, counts) == 0
    return np.frombuffer(counts, dtype=np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("scheme", [0, 1, 3])
def test_grid_orders_visit_every_tile_once(hip, scheme, grid):
    counts = visits(hip, scheme, grid)
    assert (counts[:-1] == 1).all() and counts[-1] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("workgroups", [8, 16, 24, 768])
def test_the_banded_walk_visits_every_tile_once(hip, grid, workgroups):
    counts = visits(hip, 2, grid, workgroups)
    assert (counts[:-1] == 1).all() and counts[-1] == 0


@pytest.mark.gpu
def test_walks_the_kernels_never_launch_are_refused(hip):
    counts = (ctypes.c_uint * 9)()
    for workgroups in (0, 4, 12):
        assert hip.ftc_tile_visit_probe_v1(2, 2, 2, 2, workgroups, counts) != 0
    assert hip.ftc_tile_visit_probe_v1(4, 2, 2, 2, 8, counts) != 0
