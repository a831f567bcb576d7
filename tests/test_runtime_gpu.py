"""The runtime's graph queue on the device: a merged member launch (scoped_runtime.h:401-413) runs one
kernel whose members' effects are committed as their union, and a refused one runs nothing.

The task body is the library's iota kernel (x[i] = scale * i), so the expected contents are exact.
"""
import ctypes
from ctypes import c_size_t, c_void_p

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Kernel(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("function", c_void_p),
                ("grid", ctypes.c_uint32 * 3), ("block", ctypes.c_uint32 * 3), ("shared_bytes", c_size_t),
                ("arguments", c_void_p), ("argument_count", c_size_t)]


class Member(ctypes.Structure):
    _fields_ = [("bindings", c_void_p), ("binding_count", c_size_t)]


def iota_launch(hip, device_ptr, n, scale):
    hip.ftc_iota_kernel.restype = c_void_p
    x = c_void_p(device_ptr)
    nn = ctypes.c_longlong(n)
    sc = ctypes.c_double(scale)
    keep = (x, nn, sc)
    argv = (c_void_p * 3)(*(ctypes.addressof(v) for v in keep))
    sizes = (c_size_t * 3)(8, 8, 8)
    k = Kernel(1, 0, c_void_p(hip.ftc_iota_kernel()), (ctypes.c_uint32 * 3)((n + 255) // 256, 1, 1),
               (ctypes.c_uint32 * 3)(256, 1, 1), 0, ctypes.cast(argv, c_void_p), 3)
    return k, sizes, (keep, argv)


def members_of(scope, groups):
    keep, out = [], []
    for uses in groups:
        binds = []
        for h, kw in uses:
            acc, kp = scope.make_access(**kw)
            keep.append(kp)
            binds.append(scope.PlanBinding(h, acc))
        arr = (scope.PlanBinding * len(binds))(*binds)
        keep.append(arr)
        out.append(Member(ctypes.cast(arr, c_void_p), len(binds)))
    return (Member * len(out))(*out), keep


def test_member_launch_commits_the_union_of_its_members(scope, hip):
    n = 4096
    a = np.full(n, -1.0)
    with scope.Domain(hip) as dom:
        h = dom.bind(a)
        dev = c_void_p()
        scope.check(hip, hip.fort_scope_device_reserve_v1(dom.ctx, h, ctypes.byref(dev)))
        k, sizes, keep = iota_launch(hip, dev.value, n, 0.5)
        half = n // 2
        ms, keep2 = members_of(scope, [[(h, {"overwrites": [((0,), (half,))]})],
                                       [(h, {"overwrites": [((half,), (n,))]})]])
        scope.check(hip, hip.fort_scope_graph_enqueue_members_v1(dom.ctx, 11, ctypes.byref(k), sizes, ms, 2))
        scope.check(hip, hip.fort_scope_graph_flush_v1(dom.ctx))
        dom.observe(h)
    assert np.array_equal(a, 0.5 * np.arange(n))


def test_refused_member_launch_runs_nothing(scope, hip):
    n = 1024
    a = np.full(n, 3.0)
    with scope.Domain(hip) as dom:
        h = dom.bind(a)
        dev = c_void_p()
        scope.check(hip, hip.fort_scope_device_reserve_v1(dom.ctx, h, ctypes.byref(dev)))
        k, sizes, keep = iota_launch(hip, dev.value, n, 2.0)
        ms, keep2 = members_of(scope, [[(h, {"writes": [((0,), (600,))]})], [(h, {"reads": [((500,), (n,))]})]])
        assert hip.fort_scope_graph_enqueue_members_v1(dom.ctx, 12, ctypes.byref(k), sizes, ms, 2) == scope.BOUNDARY
        scope.check(hip, hip.fort_scope_graph_flush_v1(dom.ctx))
        dom.observe(h)
    assert np.array_equal(a, np.full(n, 3.0))
