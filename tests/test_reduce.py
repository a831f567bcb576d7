"""Fixed-order reductions: the oracle is pinned on the CPU (order-free operations equal the serial result,
integer-valued sums are exact), then the device reduction is compared with it bit for bit (same order)."""
import numpy as np
import pytest

from conftest import bits_equal

SUM, SUMSQ, MAX, MIN, MAXABS = range(5)


@pytest.mark.parametrize("shape,lo,hi", [((20, 10, 6), (1, 1, 1), (19, 9, 5)),
                                         ((70, 40, 30), (0, 0, 0), (70, 40, 30))])
def test_oracle_order_free_results_equal_serial(oracle, shape, lo, hi):
    rng = np.random.default_rng(1)
    x = np.asfortranarray(rng.uniform(-3, 3, shape))
    box = x[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    assert oracle.fixed_order_reduce(x, lo, hi, MAX, -np.inf) == box.max()
    assert oracle.fixed_order_reduce(x, lo, hi, MIN, np.inf) == box.min()
    assert oracle.fixed_order_reduce(x, lo, hi, MAXABS, 0.0) == np.abs(box).max()
    ints = np.asfortranarray(rng.integers(-1000, 1000, shape).astype(np.float64))
    ibox = ints[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    assert oracle.fixed_order_reduce(ints, lo, hi, SUM, 5.0) == float(ibox.sum()) + 5.0
    assert oracle.fixed_order_reduce(ints, lo, hi, SUMSQ, 0.0) == float((ibox * ibox).sum())


def test_oracle_real_sum_is_close_to_serial(oracle):
    rng = np.random.default_rng(2)
    x = np.asfortranarray(rng.uniform(0, 1, (100, 100, 20)))
    got = oracle.fixed_order_reduce(x, (0, 0, 0), x.shape, SUM, 0.0)
    assert abs(got - x.sum()) <= 1e-10 * x.sum()


@pytest.mark.gpu
@pytest.mark.parametrize("op", [SUM, SUMSQ, MAX, MIN, MAXABS])
@pytest.mark.parametrize("shape,lo,hi", [((20, 10, 6), (1, 1, 1), (19, 9, 5)),
                                         ((130, 100, 40), (1, 1, 1), (129, 99, 39)),
                                         ((66, 66, 66), (0, 0, 0), (66, 66, 66))])
def test_device_reduction_is_bit_exact(ftc, scope, hip, oracle, op, shape, lo, hi):
    import ctypes

    rng = np.random.default_rng(op + sum(shape))
    x = np.asfortranarray(rng.uniform(-2, 2, shape))
    initial = 0.5 if op in (SUM, SUMSQ) else (-np.inf if op == MAX else (np.inf if op == MIN else 0.0))
    value = ctypes.c_double(initial)
    lo3 = (ctypes.c_longlong * 3)(*lo)
    hi3 = (ctypes.c_longlong * 3)(*hi)
    with scope.Domain(hip) as dom:
        h = dom.bind(x)
        scope.check(hip, hip.cpp_fort_shared_fixed_order_reduction(dom.ctx, 1, h, ftc.ip(op), lo3, hi3,
                                                                   ctypes.byref(value)))
    want = oracle.fixed_order_reduce(x, lo, hi, op, initial)
    assert bits_equal(np.array([value.value]), np.array([want]))
