"""Fixed-order reductions: the oracle is pinned on the CPU (order-free operations equal the serial result,
integer-valued sums are exact, and two known answers fix the order of a real sum), then the device reduction
is compared with it bit for bit and with the same known answers."""
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


B = 2.0 ** 53

# Order-sensitive known answers for a real sum (compiler/resident/reductions.py:18-31, SLOTS = 65536,
# FINAL = 256, DIRECT = 65536 at :66-72). The values sit where the spacing of doubles is 2 or 4, so each
# addition of an odd integer rounds a tie to the even neighbour, and the result records the order of the
# additions. Derived by hand:
#
# DIRECT (1024 iterations; final value f takes iterations f, f+256, f+512, f+768; the host folds the 256
# values in index order into the initial value 1):
#   x[0] = 2, x[1] = 2, x[512] = 3, x[768] = -2^54
#   f = 0: 2 + 0 + 3 = 5; 5 + (-2^54) = -(2^54 - 5), a tie between -(2^54 - 4) and -(2^54 - 6): -(2^54 - 4)
#   f = 1: 2
#   host: 1 + (-(2^54 - 4)) = -(2^54 - 5) -> -(2^54 - 4); + 2 = -(2^54 - 6), exact.  Result -2^54 + 6.
#   (A serial sum gives -2^54 + 8, contiguous chunks -2^54 + 8, a reversed fold -2^54 + 8.)
#
# SLOTS (131072 iterations; slot s takes s and s + 65536; final value f takes slots f, f+256, ...):
#   x[1] = 3, x[256] = 1, x[257] = 3, x[65537] = 2^53
#   slot 1: 3 + 2^53 = 2^53 + 3, a tie: 2^53 + 4; slot 256: 1; slot 257: 3
#   f = 0: 0 + 1 = 1; f = 1: (2^53 + 4) + 3 = 2^53 + 7, a tie: 2^53 + 8
#   host: 1 + 1 = 2; 2 + 2^53 + 8 = 2^53 + 10, exact.  Result 2^53 + 10.
#   (A serial sum gives 2^53 + 8, and so do contiguous slots, contiguous folds and a reversed fold.)
ORDER_KATS = [
    ((16, 16, 4), {0: 2.0, 1: 2.0, 512: 3.0, 768: -2.0 * B}, 1.0, -2.0 * B + 6.0),
    ((64, 64, 32), {1: 3.0, 256: 1.0, 257: 3.0, 65537: B}, 1.0, B + 10.0),
]


def order_kat_field(shape, values):
    x = np.zeros(int(np.prod(shape)))
    for t, v in values.items():
        x[t] = v
    return np.asfortranarray(x.reshape(shape, order="F"))


@pytest.mark.parametrize("shape,values,initial,want", ORDER_KATS)
def test_oracle_real_sum_follows_the_strided_order(oracle, shape, values, initial, want):
    x = order_kat_field(shape, values)
    got = oracle.fixed_order_reduce(x, (0, 0, 0), shape, SUM, initial)
    assert got == want, (got - want)
    serial = initial
    for v in x.ravel(order="F"):
        serial += v
    assert serial != want


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


@pytest.mark.gpu
@pytest.mark.parametrize("shape,values,initial,want", ORDER_KATS)
def test_device_real_sum_gives_the_order_known_answers(ftc, scope, hip, shape, values, initial, want):
    import ctypes

    x = order_kat_field(shape, values)
    value = ctypes.c_double(initial)
    lo3 = (ctypes.c_longlong * 3)(0, 0, 0)
    hi3 = (ctypes.c_longlong * 3)(*shape)
    with scope.Domain(hip) as dom:
        h = dom.bind(x)
        scope.check(hip, hip.cpp_fort_shared_fixed_order_reduction(dom.ctx, 1, h, ftc.ip(SUM), lo3, hi3,
                                                                   ctypes.byref(value)))
    assert value.value == want, value.value - want
