"""The fixed-order reduction entry that allocates nothing (ftc_fixed_order_reduction_static_v1): its slot
and fold values live in the library's static device scratch. Called through the C ABI on device pointers
of arrays the runtime keeps resident, it must give bit for bit what cpp_fort_shared_fixed_order_reduction
gives on the same box, on consecutive calls that reuse the scratch, the later ones smaller."""
import ctypes
from ctypes import byref, c_int, c_void_p

import numpy as np
import pytest

from conftest import bits_equal

SUM, SUMSQ, MAX, MIN, MAXABS = range(5)
INITIAL = {SUM: 0.5, SUMSQ: 0.25, MAX: -np.inf, MIN: np.inf, MAXABS: 0.0}


def reference_entry(ftc, scope, hip, x, lo, hi, op, initial):
    value = ctypes.c_double(initial)
    with scope.Domain(hip) as dom:
        h = dom.bind(x)
        scope.check(hip, hip.cpp_fort_shared_fixed_order_reduction(
            dom.ctx, 1, h, ftc.ip(op), (ctypes.c_longlong * 3)(*lo), (ctypes.c_longlong * 3)(*hi), byref(value)))
    return value.value


def static_entry(scope, hip, dom, x, lo, hi, op, initial):
    """Runs the static entry on the device mirrors of x (read) and of a one-value result (written), on the
    context stream, and returns the result brought back by observation."""
    result = np.zeros(1)
    hx, hr = dom.bind(x), dom.bind(result, initialized=False)
    access_x, keep_x = scope.make_access(scope.READ_ALL)
    access_r, keep_r = scope.make_access(0, (), [((0,), (1,))], [((0,), (1,))])
    plan = (scope.PlanBinding * 2)(scope.PlanBinding(hx, access_x), scope.PlanBinding(hr, access_r))
    devices = (c_void_p * 2)()
    scope.check(hip, hip.fort_scope_access_begin_v1(dom.ctx, plan, 2, 1, devices))
    prev, stream = c_int(0), c_void_p()
    scope.check(hip, hip.fort_scope_gpu_enter(dom.ctx, byref(prev), byref(stream)))
    ld1, ld2 = x.shape[0], x.shape[0] * x.shape[1]
    start = devices[0] + 8 * (lo[0] + lo[1] * ld1 + lo[2] * ld2)
    status = hip.ftc_fixed_order_reduction_static_v1(start, hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], ld1, ld2,
                                                      op, initial, devices[1], stream)
    scope.check(hip, hip.fort_scope_gpu_leave(dom.ctx, prev.value))
    scope.check(hip, status)
    scope.check(hip, hip.fort_scope_device_end(dom.ctx, hx))
    scope.check(hip, hip.fort_scope_device_end(dom.ctx, hr))
    dom.observe(hr)
    return float(result[0])


# Two-level path (more than 65,536 iterations), then the direct path, then an empty box.
CASES = [((130, 100, 40), (1, 1, 1), (129, 99, 39)),
         ((20, 10, 6), (1, 1, 1), (19, 9, 5)),
         ((20, 10, 6), (3, 2, 1), (3, 9, 5))]


@pytest.mark.gpu
@pytest.mark.parametrize("op", [SUM, SUMSQ, MAX, MIN, MAXABS])
def test_static_scratch_reduction_equals_the_entry_on_consecutive_calls(ftc, scope, hip, op):
    print(f"static reduction op={op}", flush=True)
    rng = np.random.default_rng(40 + op)
    fields = [np.asfortranarray(rng.uniform(-2, 2, shape)) for shape, _, _ in CASES]
    wants = [reference_entry(ftc, scope, hip, x, lo, hi, op, INITIAL[op]) for x, (_, lo, hi) in zip(fields, CASES)]
    with scope.Domain(hip) as dom:
        for x, (shape, lo, hi), want in zip(fields, CASES, wants):
            got = static_entry(scope, hip, dom, x, lo, hi, op, INITIAL[op])
            assert bits_equal(np.array([got]), np.array([want])), (shape, lo, hi, got, want)


@pytest.mark.gpu
def test_static_scratch_reduction_keeps_a_negative_zero_for_an_empty_box(scope, hip):
    x = np.asfortranarray(np.ones((4, 3, 2)))
    with scope.Domain(hip) as dom:
        got = static_entry(scope, hip, dom, x, (0, 0, 0), (4, 0, 2), SUM, -0.0)
    assert np.signbit(got) and got == 0.0


def test_static_scratch_reduction_rejects_bad_arguments_before_any_device_call(ftc):
    lib = ftc.load()
    result = ctypes.c_double()
    x = ctypes.c_double()
    cases = [(dict(x=None), b"null device pointer"),
             (dict(op=7), b"unknown reduction"),
             (dict(n1=-1), b"negative extent"),
             (dict(ld1=1), b"leading dimensions below the box extents"),
             (dict(n1=1 << 40, n2=1 << 40, ld1=1 << 40, ld2=1 << 41), b"reduction box beyond 2^62 elements")]
    for change, message in cases:
        a = dict(x=ctypes.addressof(x), n1=4, n2=3, n3=2, ld1=4, ld2=12, op=SUM, initial=0.0,
                 result=ctypes.addressof(result), stream=None)
        a.update(change)
        status = lib.ftc_fixed_order_reduction_static_v1(a["x"], a["n1"], a["n2"], a["n3"], a["ld1"], a["ld2"],
                                                         a["op"], a["initial"], a["result"], a["stream"])
        assert status != 0 and lib.fort_scope_error() == message, (change, status, lib.fort_scope_error())
