"""The runtime on the device: a merged member launch (scoped_runtime.h:401-413) runs one kernel whose
members' effects are committed as their union, and a refused one runs nothing; original host work and the
identity exchange (scoped_runtime.h:198-211, 234-246) with device-resident values.

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


def overwrite(lo, hi):
    """A rank-1 access that writes all of [lo, hi) (a must-overwrite section is also a may-write one)."""
    return {"writes": [((lo,), (hi,))], "overwrites": [((lo,), (hi,))]}


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
        ms, keep2 = members_of(scope, [[(h, overwrite(0, half))],
                                       [(h, overwrite(half, n))]])
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


def test_original_host_read_publishes_only_device_defined_cells(scope, hip):
    # Original CPU work over a partly defined array (compiler/tests/test_original_host_access.py): the
    # device-defined cells reach the host, the holes stay as they were and stay undefined.
    n = 8
    a = np.full(n, -99.0)
    with scope.Domain(hip) as dom:
        h = dom.bind(a, initialized=False)
        dev = c_void_p()
        scope.check(hip, hip.fort_scope_device_reserve_v1(dom.ctx, h, ctypes.byref(dev)))
        k, sizes, keep = iota_launch(hip, dev.value + 2 * 8, 4, 7.0)  # cells 2..5 = 0, 7, 14, 21
        ms, keep2 = members_of(scope, [[(h, overwrite(2, 6))]])
        scope.check(hip, hip.fort_scope_graph_enqueue_members_v1(dom.ctx, 13, ctypes.byref(k), sizes, ms, 1))
        scope.check(hip, hip.fort_scope_graph_flush_v1(dom.ctx))
        acc, kp = scope.make_access(scope.READ_ALL)
        b = (scope.PlanBinding * 1)(scope.PlanBinding(h, acc))
        scope.check(hip, hip.fort_scope_original_host_begin_v1(dom.ctx, b, 1))
        assert list(a) == [-99, -99, 0, 7, 14, 21, -99, -99]
        scope.check(hip, hip.fort_scope_host_end(dom.ctx, h))
        ptr = c_void_p()
        assert hip.fort_scope_device_begin(dom.ctx, h, ctypes.byref(acc), ctypes.byref(ptr)) == scope.UNINITIALIZED


def test_identity_swap_keeps_the_device_copies_with_their_allocations(scope, hip):
    n = 1024
    a, b = np.zeros(n), np.zeros(n)
    with scope.Domain(hip) as dom:
        ha, hb = dom.bind(a, identity=1), dom.bind(b, identity=2)
        for h, scale in ((ha, 1.0), (hb, 2.0)):
            dev = c_void_p()
            scope.check(hip, hip.fort_scope_device_reserve_v1(dom.ctx, h, ctypes.byref(dev)))
            k, sizes, keep = iota_launch(hip, dev.value, n, scale)
            ms, keep2 = members_of(scope, [[(h, overwrite(0, n))]])
            scope.check(hip, hip.fort_scope_graph_enqueue_members_v1(dom.ctx, 14, ctypes.byref(k), sizes, ms, 1))
        scope.check(hip, hip.fort_scope_graph_flush_v1(dom.ctx))
        layouts = []
        for arr in (b, a):  # after the exchange identity 1 names b's allocation and identity 2 a's
            ext, low = (c_size_t * 1)(n), (ctypes.c_int64 * 1)(1)
            layouts.append((scope.Layout(1, scope.REAL64, 8, arr.ctypes.data, ext, low, 0), ext, low))
        cur1, cur2 = ctypes.c_uint64(), ctypes.c_uint64()
        scope.check(hip, hip.fort_scope_identity_swap_v1(dom.ctx, ha, hb, ctypes.byref(layouts[0][0]),
                                                         ctypes.byref(layouts[1][0]), ctypes.byref(cur1),
                                                         ctypes.byref(cur2)))
        assert (cur1.value, cur2.value) == (hb, ha)
        dom.observe(cur1.value, cur2.value)
    assert np.array_equal(a, np.arange(n) * 1.0) and np.array_equal(b, np.arange(n) * 2.0)


class GraphTask(ctypes.Structure):
    _fields_ = [("kernel", Kernel), ("bindings", c_void_p), ("binding_count", c_size_t)]


def test_group_submit_replays_with_updated_arguments(scope, hip):
    # graph_submit_v1 (scoped_runtime.h:349-360): resources acquired by the caller, one chain graph per
    # group key, every replay with the new argument values.
    n = 2048
    a = np.zeros(n)
    with scope.Domain(hip) as dom:
        h = dom.bind(a, initialized=False)
        for scale in (1.0, 3.0, 5.0):
            acc, kp = scope.make_access(0, (), [((0,), (n,))], [((0,), (n,))])
            b = (scope.PlanBinding * 1)(scope.PlanBinding(h, acc))
            devs = (c_void_p * 1)()
            scope.check(hip, hip.fort_scope_access_begin_v1(dom.ctx, b, 1, 1, devs))
            k, sizes, keep = iota_launch(hip, devs[0], n, scale)
            scope.check(hip, hip.fort_scope_graph_submit_v1(dom.ctx, 77, ctypes.byref(k), 1))
            scope.check(hip, hip.fort_scope_device_end(dom.ctx, h))
            dom.observe(h)
            assert np.array_equal(a, scale * np.arange(n)), scale
        g = dom.graph_stats()
        assert g.instantiations >= 1 and g.replays >= 2


def test_group_execution_uploads_only_external_requirements(scope, hip):
    n = 1024
    src = np.arange(n, dtype=np.float64)
    mid, out = np.zeros(n), np.zeros(n)
    with scope.Domain(hip) as dom:
        hs, hm, ho = dom.bind(src), dom.bind(mid, initialized=False), dom.bind(out, initialized=False)
        ptrs = []
        for h in (hm, ho):
            dev = c_void_p()
            scope.check(hip, hip.fort_scope_device_reserve_v1(dom.ctx, h, ctypes.byref(dev)))
            ptrs.append(dev.value)
        keep = []

        def task(uses, ptr, scale):
            arr = (scope.PlanBinding * len(uses))(*[scope.PlanBinding(h, acc) for h, acc in uses])
            k, sizes, kk = iota_launch(hip, ptr, n, scale)
            keep.extend([arr, kk, k])
            return GraphTask(k, ctypes.cast(arr, c_void_p), len(uses))

        r_src, k1 = scope.make_access(scope.READ_ALL)
        w_mid, k2 = scope.make_access(0, (), [((0,), (n,))], [((0,), (n,))])
        r_mid, k3 = scope.make_access(scope.READ_ALL)
        w_out, k4 = scope.make_access(0, (), [((0,), (n,))], [((0,), (n,))])
        tasks = (GraphTask * 2)(task([(hs, r_src), (hm, w_mid)], ptrs[0], 2.0), task([(hm, r_mid), (ho, w_out)],
                                                                                      ptrs[1], 4.0))
        before = dom.stats().upload_bytes
        scope.check(hip, hip.fort_scope_graph_execute_v1(dom.ctx, 5, tasks, 2))
        # The intermediate is defined inside the group: only the source went up.
        assert dom.stats().upload_bytes - before == n * 8
        dom.observe(hm, ho)
    assert np.array_equal(mid, 2.0 * np.arange(n)) and np.array_equal(out, 4.0 * np.arange(n))
