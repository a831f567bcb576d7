"""The fort_scope runtime on its CPU reference backend (libftc_cpu.so: "device" memory is host memory,
device kernels are refused with RESOURCE), as the reference tests its runtime without a GPU
(compiler/tests/test_scoped_runtime.py, test_scoped_domains.py, test_scoped_scratch.py).

Covers registration and layouts, exact section coherence (only missing sections move), definition
checks, aliasing, LRU eviction with publication of dirty mirrors, scratch leases, the external-provider
transaction, poisoning and abandonment.
"""
import ctypes
from ctypes import byref, c_int, c_size_t, c_uint64, c_void_p

import numpy as np
import pytest


def new_domain(scope, lib, budget=None):
    return scope.Domain(lib, budget=budget)


def binding(scope, handle, flags=0, reads=(), writes=(), overwrites=()):
    access, keep = scope.make_access(flags, reads, writes, overwrites)
    return scope.PlanBinding(handle, access), keep


def begin_device(scope, lib, dom, bindings):
    arr = (scope.PlanBinding * len(bindings))(*[b for b, _ in bindings])
    devs = (c_void_p * len(bindings))()
    status = lib.fort_scope_access_begin_v1(dom.ctx, arr, len(bindings), 1, devs)
    return status, [devs[i] for i in range(len(bindings))]


def dev_view(ptr, shape):
    n = int(np.prod(shape))
    buf = (ctypes.c_double * n).from_address(ptr)
    return np.frombuffer(buf, dtype=np.float64).reshape(shape, order="F")


def test_register_and_layout(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.zeros((5, 4, 3), order="F")
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        layout = scope.Layout()
        scope.check(lib, lib.fort_scope_layout_get(dom.ctx, h, byref(layout)))
        assert (layout.rank, layout.type, layout.element_bytes) == (3, scope.REAL64, 8)
        assert [layout.extents[i] for i in range(3)] == [5, 4, 3]
        assert layout.host == a.ctypes.data
        # The same identity registered again returns the same handle.
        layout2 = scope.Layout(3, scope.REAL64, 8, a.ctypes.data, layout.extents, layout.lower_bounds, 0)
        again = c_uint64()
        identity = 1
        scope.check(lib, lib.fort_scope_register(dom.ctx, identity, 0, byref(layout2), 1, byref(again)))
        assert again.value == h


def test_device_writes_reach_the_host_only_where_written(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(6 * 5 * 4, dtype=np.float64).reshape((6, 5, 4), order="F").copy(order="F")
    original = a.copy(order="F")
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        box = ((1, 1, 1), (4, 3, 2))
        b = binding(scope, h, reads=[box], writes=[box], overwrites=[box])
        st, devs = begin_device(scope, lib, dom, [b])
        scope.check(lib, st)
        s = dom.stats()
        # Only the read box went up.
        assert s.upload_bytes == 3 * 2 * 1 * 8
        view = dev_view(devs[0], a.shape)
        view[1:4, 1:3, 1:2] = -view[1:4, 1:3, 1:2]
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        assert np.array_equal(a, original)  # the host copy is untouched until observation
        dom.observe(h)
        assert dom.stats().download_bytes == 3 * 2 * 1 * 8
    expect = original.copy(order="F")
    expect[1:4, 1:3, 1:2] *= -1
    assert np.array_equal(a, expect)


def test_a_second_read_of_a_current_section_moves_nothing(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.ones((8, 8), order="F")
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        for _ in range(2):
            b = binding(scope, h, flags=scope.READ_ALL)
            st, _ = begin_device(scope, lib, dom, [b])
            scope.check(lib, st)
            scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        assert dom.stats().uploads >= 1
        assert dom.stats().upload_bytes == 64 * 8


def test_reading_undefined_values_is_refused(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.zeros((4, 4), order="F")
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL)])
        assert st == scope.UNINITIALIZED
        # An overwrite defines the section; a later read of it is accepted.
        half = ((0, 0), (4, 2))
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, writes=[half], overwrites=[half])])
        scope.check(lib, st)
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, reads=[((0, 0), (4, 2))])])
        scope.check(lib, st)
        scope.check(lib, lib.fort_scope_cancel_access(dom.ctx, h))
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, reads=[((0, 0), (4, 3))])])
        assert st == scope.UNINITIALIZED


def test_overwritten_sections_are_not_copied_and_must_be_writable(scope, cpu_runtime):
    # compiler/tests/test_scoped_runtime.py:264-291 (an overwrite skips input) and the may-write rule of
    # compiler/runtime/scoped_runtime.cu:816-817, here on defined values that are not current on the device.
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        # Writes over [1, 7), of which [2, 6) is overwritten: only the preserved cells 1 and 6 go up.
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, writes=[((1,), (7,))], overwrites=[((2,), (6,))])])
        scope.check(lib, st)
        assert (dom.stats().uploads, dom.stats().upload_bytes) == (2, 2 * 8)
        dev_view(devs[0], (8,))[1:7] = -1.0
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        every = scope.WRITE_ALL | scope.OVERWRITE_ALL
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, flags=every)])
        scope.check(lib, st)
        assert dom.stats().upload_bytes == 2 * 8  # nothing more
        dev_view(devs[0], (8,))[:] = 5.0
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        for bad in (dict(overwrites=[((0,), (2,))]), dict(writes=[((0,), (3,))], overwrites=[((2,), (5,))]),
                    dict(flags=scope.OVERWRITE_ALL)):
            st, _ = begin_device(scope, lib, dom, [binding(scope, h, **bad)])
            assert st == scope.ARGUMENT
        # Empty sections name nothing: an empty overwrite needs no write.
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, reads=[((0,), (8,))], overwrites=[((3,), (3,))])])
        scope.check(lib, st)
        scope.check(lib, lib.fort_scope_cancel_access(dom.ctx, h))
        dom.observe(h)
    assert np.array_equal(a, np.full(8, 5.0))


def test_repeated_buffers_and_bad_sections_are_rejected(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.zeros((4, 4), order="F")
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL),
                                               binding(scope, h, flags=scope.READ_ALL)])
        assert st == scope.ALIAS
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, reads=[((0, 0), (5, 1))])])
        assert st == scope.ARGUMENT


def test_budget_pressure_evicts_the_least_recently_used_and_publishes_it(scope, cpu_runtime):
    lib = cpu_runtime
    n = 1000
    a, b, c = (np.full((n,), float(v), order="F") for v in (1, 2, 3))
    with new_domain(scope, lib, budget=2 * n * 8 + 16) as dom:
        ha, hb, hc = dom.bind(a), dom.bind(b), dom.bind(c)
        for h in (ha, hb):
            every = scope.READ_ALL | scope.WRITE_ALL | scope.OVERWRITE_ALL
            st, devs = begin_device(scope, lib, dom, [binding(scope, h, flags=every)])
            scope.check(lib, st)
            dev_view(devs[0], (n,))[:] += 10.0
            scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        st, _ = begin_device(scope, lib, dom, [binding(scope, hc, flags=scope.READ_ALL)])
        scope.check(lib, st)
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, hc))
        res = scope.ResidencyStats()
        scope.check(lib, lib.fort_scope_residency_stats_get_v1(dom.ctx, byref(res)))
        assert res.evictions == 1 and res.dirty_evictions == 1
        # The evicted buffer (a, least recently used) was published before its mirror went away.
        assert np.all(a == 11.0)
        assert np.all(b == 2.0)
    assert np.all(b == 12.0) and np.all(c == 3.0)


def test_a_footprint_larger_than_the_budget_fails_before_any_change(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.zeros(100, order="F")
    with new_domain(scope, lib, budget=400) as dom:
        h = dom.bind(a)
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL)])
        assert st == scope.RESOURCE
        assert dom.stats().allocations == 0


def test_device_kernels_are_refused_by_the_cpu_backend(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.zeros(16, order="F")
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        access, keep = scope.make_access(scope.READ_ALL)
        bindings = (scope.PlanBinding * 1)(scope.PlanBinding(h, access))

        class Kernel(ctypes.Structure):
            _fields_ = [("version", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("function", c_void_p),
                        ("grid", ctypes.c_uint32 * 3), ("block", ctypes.c_uint32 * 3), ("shared_bytes", c_size_t),
                        ("arguments", c_void_p), ("argument_count", c_size_t)]

        class Task(ctypes.Structure):
            _fields_ = [("kernel", Kernel), ("bindings", c_void_p), ("binding_count", c_size_t)]

        arg = ctypes.c_double(1.0)
        argv = (c_void_p * 1)(ctypes.addressof(arg))
        sizes = (c_size_t * 1)(8)
        k = Kernel(1, 0, c_void_p(1), (ctypes.c_uint32 * 3)(1, 1, 1), (ctypes.c_uint32 * 3)(64, 1, 1), 0,
                   ctypes.cast(argv, c_void_p), 1)
        t = Task(k, ctypes.cast(bindings, c_void_p), 1)
        st = lib.fort_scope_graph_enqueue_v1(dom.ctx, 1, ctypes.addressof(t), sizes)
        assert st == scope.RESOURCE
        assert dom.stats().allocations == 0


def test_member_launches_are_checked_on_their_actual_sections(scope, cpu_runtime):
    # scoped_runtime.h:401-413: members that write a section another member uses are refused with BOUNDARY
    # (nothing queued, context still usable); disjoint members are accepted, which the CPU backend then
    # refuses with RESOURCE because it runs no device kernels.
    lib = cpu_runtime
    a = np.zeros(16, order="F")
    b = np.zeros(16, order="F")

    class Kernel(ctypes.Structure):
        _fields_ = [("version", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("function", c_void_p),
                    ("grid", ctypes.c_uint32 * 3), ("block", ctypes.c_uint32 * 3), ("shared_bytes", c_size_t),
                    ("arguments", c_void_p), ("argument_count", c_size_t)]

    class Member(ctypes.Structure):
        _fields_ = [("bindings", c_void_p), ("binding_count", c_size_t)]

    arg = ctypes.c_double(1.0)
    argv = (c_void_p * 1)(ctypes.addressof(arg))
    sizes = (c_size_t * 1)(8)
    k = Kernel(1, 0, c_void_p(1), (ctypes.c_uint32 * 3)(1, 1, 1), (ctypes.c_uint32 * 3)(64, 1, 1), 0,
               ctypes.cast(argv, c_void_p), 1)
    with new_domain(scope, lib) as dom:
        ha, hb = dom.bind(a), dom.bind(b)
        keep = []

        def member(*uses):
            binds = []
            for h, kw in uses:
                acc, kp = scope.make_access(**kw)
                keep.append(kp)
                binds.append(scope.PlanBinding(h, acc))
            arr = (scope.PlanBinding * len(binds))(*binds)
            keep.append(arr)
            return Member(ctypes.cast(arr, c_void_p), len(binds))

        def launch(*members):
            ms = (Member * len(members))(*members)
            return lib.fort_scope_graph_enqueue_members_v1(dom.ctx, 7, ctypes.byref(k), sizes, ms, len(members))

        w = lambda lo, hi: {"writes": [((lo,), (hi,))]}
        r = lambda lo, hi: {"reads": [((lo,), (hi,))]}
        assert launch(member((ha, w(0, 8))), member((ha, w(8, 16)))) == scope.RESOURCE
        assert launch(member((ha, w(0, 9))), member((ha, r(8, 16)))) == scope.BOUNDARY
        assert launch(member((ha, r(0, 16))), member((ha, r(0, 16)), (hb, w(0, 16)))) == scope.RESOURCE
        whole_b = {"writes": [((0,), (16,))], "overwrites": [((0,), (16,))]}
        assert launch(member((ha, r(0, 4))), member((hb, whole_b), (ha, w(4, 16)))) == scope.RESOURCE
        assert launch(member((ha, {"flags": scope.WRITE_ALL})), member((ha, r(15, 16)))) == scope.BOUNDARY
        assert launch(member((ha, r(0, 4)), (ha, r(4, 8)))) == scope.ALIAS
        assert dom.stats().allocations == 0
        dom.observe(ha)  # a refused launch does not poison the context


def test_scratch_lease(scope, cpu_runtime):
    lib = cpu_runtime
    with new_domain(scope, lib) as dom:
        lease = scope.ScratchLease()
        scope.check(lib, lib.fort_scope_scratch_acquire_v1(dom.ctx, 4096, byref(lease)))
        assert lease.device and lease.capacity >= 4096
        again = scope.ScratchLease()
        assert lib.fort_scope_scratch_acquire_v1(dom.ctx, 16, byref(again)) == scope.STATE
        scope.check(lib, lib.fort_scope_scratch_release_v1(dom.ctx, lease.token))
        scope.check(lib, lib.fort_scope_scratch_acquire_v1(dom.ctx, 1024, byref(again)))
        assert again.device == lease.device  # reused arena
        scope.check(lib, lib.fort_scope_scratch_release_v1(dom.ctx, again.token))


def test_external_provider_transaction(scope, cpu_runtime):
    lib = cpu_runtime
    rhs = np.arange(12, dtype=np.float64).reshape((3, 4), order="F").copy(order="F")
    phi = np.zeros((3, 4), order="F")
    seen = {}

    @scope.EXTERNAL_WORKER
    def worker(user, views, count, stream, started):
        started[0] = 1
        seen["count"] = count
        seen["extents"] = [views[0].extents[d] for d in range(2)]
        seen["strides"] = [views[0].strides[d] for d in range(2)]
        src = dev_view(views[1].device, (3, 4))
        dst = dev_view(views[0].device, (3, 4))
        dst[:] = 2.0 * src
        return 0

    with new_domain(scope, lib) as dom:
        hp, hr = dom.bind(phi, initialized=False), dom.bind(rhs)
        a0, k0 = scope.make_access(scope.WRITE_ALL | scope.OVERWRITE_ALL)
        a1, k1 = scope.make_access(scope.READ_ALL)
        eb = (scope.ExternalBinding * 2)(scope.ExternalBinding(hp, scope.REAL64, 2, 8, a0),
                                         scope.ExternalBinding(hr, scope.REAL64, 2, 8, a1))
        call = scope.ExternalCall(1, 0, eb, 2, 0, 0)
        report = scope.ExternalReport()
        scope.check(lib, lib.fort_scope_external_execute_v1(dom.ctx, byref(call), worker, None, byref(report)))
        assert report.started == 1 and report.committed == 1
        dom.observe(hp)
    assert seen == {"count": 2, "extents": [3, 4], "strides": [8, 24]}
    assert np.array_equal(phi, 2.0 * rhs)


def test_a_provider_that_fails_after_starting_poisons_the_context(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.zeros(4, order="F")

    @scope.EXTERNAL_WORKER
    def worker(user, views, count, stream, started):
        started[0] = 1
        return scope.EXECUTION

    dom = new_domain(scope, lib)
    h = dom.bind(a)
    acc, keep = scope.make_access(scope.READ_ALL | scope.WRITE_ALL)
    eb = (scope.ExternalBinding * 1)(scope.ExternalBinding(h, scope.REAL64, 1, 8, acc))
    call = scope.ExternalCall(1, 0, eb, 1, 0, 0)
    assert lib.fort_scope_external_execute_v1(dom.ctx, byref(call), worker, None, None) == scope.EXECUTION
    assert lib.fort_scope_wait(dom.ctx) == scope.EXECUTION
    scope.check(lib, lib.fort_scope_abandon(dom.ctx))
    scope.check(lib, lib.fort_scope_domain_abandon_v1(dom.domain))


def test_a_provider_that_declines_before_enqueue_changes_nothing(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.ones(4, order="F")

    @scope.EXTERNAL_WORKER
    def worker(user, views, count, stream, started):
        started[0] = 0
        return scope.BOUNDARY

    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        acc, keep = scope.make_access(scope.READ_ALL | scope.WRITE_ALL | scope.OVERWRITE_ALL)
        eb = (scope.ExternalBinding * 1)(scope.ExternalBinding(h, scope.REAL64, 1, 8, acc))
        call = scope.ExternalCall(1, 0, eb, 1, 0, 0)
        assert lib.fort_scope_external_execute_v1(dom.ctx, byref(call), worker, None, None) == scope.BOUNDARY
        scope.check(lib, lib.fort_scope_wait(dom.ctx))  # the context is still usable


def test_execution_error_poisons_until_abandon(scope, cpu_runtime):
    lib = cpu_runtime
    dom = new_domain(scope, lib)
    assert lib.fort_scope_execution_error(dom.ctx, b"kernel failed") == scope.EXECUTION
    assert lib.fort_scope_wait(dom.ctx) == scope.EXECUTION
    assert b"only abandon" in lib.fort_scope_error()
    scope.check(lib, lib.fort_scope_abandon(dom.ctx))
    assert lib.fort_scope_wait(dom.ctx) == scope.STALE
    scope.check(lib, lib.fort_scope_domain_abandon_v1(dom.domain))


def test_one_context_per_domain(scope, cpu_runtime):
    lib = cpu_runtime
    with new_domain(scope, lib) as dom:
        other = c_uint64()
        assert lib.fort_scope_create_in_domain_v1(dom.domain, byref(other)) == scope.STATE


def test_default_budget_honours_the_environment(scope, cpu_runtime, monkeypatch):
    lib = cpu_runtime
    monkeypatch.setenv("FORT_RESIDENT_DEVICE_BUDGET", "3M")
    out = c_size_t()
    scope.check(lib, lib.fort_scope_device_budget_default_v1(0, byref(out)))
    assert out.value == 3 << 20
    # A domain created afterwards takes it, and reservation runs without deadlock.
    a = np.zeros(16, order="F")
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        devs = (c_void_p * 1)()
        hs = (c_uint64 * 1)(h)
        scope.check(lib, lib.fort_scope_device_reserve_many_v1(dom.ctx, hs, 1, devs))
        budget, unclaimed = c_size_t(), c_size_t()
        scope.check(lib, lib.fort_scope_domain_budget_get_v1(dom.domain, byref(budget), byref(unclaimed)))
        assert budget.value == 3 << 20 and unclaimed.value == (3 << 20) - 128


def test_reservation_without_a_stated_budget(scope, cpu_runtime, monkeypatch):
    # The first reservation of a domain with no stated budget queries the device (regression: it once
    # re-entered the registry lock).
    monkeypatch.delenv("FORT_RESIDENT_DEVICE_BUDGET", raising=False)
    lib = cpu_runtime
    a = np.zeros(16, order="F")
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        devs = (c_void_p * 1)()
        hs = (c_uint64 * 1)(h)
        scope.check(lib, lib.fort_scope_device_reserve_many_v1(dom.ctx, hs, 1, devs))
        assert devs[0]


# ---------------------------------------------------------------------------------------------------------
# Original host work with may-access envelopes (compiler/tests/test_original_host_access.py) and the
# descriptor-only identity exchange (compiler/tests/test_scoped_domains.py:318-432).

def original_begin(scope, lib, dom, bindings):
    arr = (scope.PlanBinding * len(bindings))(*[b for b, _ in bindings])
    return lib.fort_scope_original_host_begin_v1(dom.ctx, arr, len(bindings))


def strict_device_read(scope, lib, dom, handle, reads=(), flags=0):
    access, _keep = scope.make_access(flags, reads, (), ())
    ptr = c_void_p()
    return lib.fort_scope_device_begin(dom.ctx, handle, byref(access), byref(ptr))


def device_write(scope, lib, dom, handle, box, values):
    st, devs = begin_device(scope, lib, dom, [binding(scope, handle, writes=[box], overwrites=[box])])
    scope.check(lib, st)
    view = dev_view(devs[0], (8,))
    view[box[0][0]:box[1][0]] = values
    scope.check(lib, lib.fort_scope_device_end(dom.ctx, handle))


def test_original_host_reads_of_unknown_values_define_nothing(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        scope.check(lib, original_begin(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL)]))
        assert a[3] == 3.0
        scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
        # A numerical read still needs definitions the original read did not give.
        assert strict_device_read(scope, lib, dom, h, flags=scope.READ_ALL) == scope.UNINITIALIZED
        s = dom.stats()
        assert s.uploads == s.downloads == 0


def test_original_host_reads_publish_only_tracked_cells_and_keep_the_device_copy(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.full(8, -99.0)
    cells = ((2,), (5,))
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        device_write(scope, lib, dom, h, cells, [14.0, 21.0, 28.0])
        scope.check(lib, original_begin(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL)]))
        assert list(a) == [-99, -99, 14, 21, 28, -99, -99, -99]
        scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
        assert strict_device_read(scope, lib, dom, h, flags=scope.READ_ALL) == scope.UNINITIALIZED
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, reads=[cells])])
        scope.check(lib, st)
        assert list(dev_view(devs[0], (8,))[2:5]) == [14, 21, 28]
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        s = dom.stats()
        assert s.upload_bytes == 0 and s.download_bytes == 24  # the read kept the device replica


def test_original_possible_writes_keep_holes_and_invalidate_device_values(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.full(8, -99.0)
    known = ((2,), (5,))
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        device_write(scope, lib, dom, h, known, [2.0, 3.0, 4.0])
        scope.check(lib, original_begin(scope, lib, dom, [binding(scope, h, flags=scope.WRITE_ALL)]))
        a[3] += 10.0
        a[0] = 17.0
        scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
        assert list(a) == [17, -99, 2, 13, 4, -99, -99, -99]
        # A possible write defines nothing new ...
        assert strict_device_read(scope, lib, dom, h, reads=[((0,), (1,))]) == scope.UNINITIALIZED
        # ... but it replaces the device copy of what it may have changed.
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, reads=[known])])
        scope.check(lib, st)
        assert list(dev_view(devs[0], (8,))[2:5]) == [2, 13, 4]
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        s = dom.stats()
        assert s.upload_bytes == s.download_bytes == 24


def test_only_proved_original_overwrites_define_values(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.full(8, -99.0)
    cells = ((1,), (3,))
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        b = binding(scope, h, flags=scope.WRITE_ALL, overwrites=[cells])
        scope.check(lib, original_begin(scope, lib, dom, [b]))
        a[1], a[2] = 7.0, 9.0
        scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
        assert strict_device_read(scope, lib, dom, h, flags=scope.READ_ALL) == scope.UNINITIALIZED
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, reads=[cells])])
        scope.check(lib, st)
        assert list(dev_view(devs[0], (8,))[1:3]) == [7, 9]
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        assert dom.stats().upload_bytes == 16 and dom.stats().download_bytes == 0


def test_a_bad_later_original_binding_moves_nothing(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.full(8, -99.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        device_write(scope, lib, dom, h, ((2,), (4,)), [7.0, 9.0])
        bad = [binding(scope, h, flags=scope.READ_ALL), binding(scope, 999999, flags=scope.READ_ALL)]
        assert original_begin(scope, lib, dom, bad) != scope.OK
        assert list(a) == [-99.0] * 8 and dom.stats().download_bytes == 0
        scope.check(lib, original_begin(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL)]))
        scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
        assert list(a) == [-99, -99, 7, 9, -99, -99, -99, -99]


def register(scope, lib, dom, array, identity, generation, lower, initialized=True):
    ext = (c_size_t * array.ndim)(*array.shape)
    low = (ctypes.c_int64 * array.ndim)(*lower)
    layout = scope.Layout(array.ndim, scope.REAL64, 8, array.ctypes.data, ext, low, generation)
    h = c_uint64()
    status = lib.fort_scope_register(dom.ctx, identity, generation, byref(layout), 1 if initialized else 0, byref(h))
    return status, h.value, (layout, ext, low)


def layout_of(scope, array, generation, lower):
    ext = (c_size_t * array.ndim)(*array.shape)
    low = (ctypes.c_int64 * array.ndim)(*lower)
    return scope.Layout(array.ndim, scope.REAL64, 8, array.ctypes.data, ext, low, generation), (ext, low)


def test_identity_swap_exchanges_only_the_identity_map(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(1.0, 5.0).reshape((2, 2), order="F")
    b = np.arange(11.0, 17.0).reshape((3, 2), order="F")
    with new_domain(scope, lib) as dom:
        st, ha, _ = register(scope, lib, dom, a, 1, 11, (-2, 3))
        scope.check(lib, st)
        st, hb, _ = register(scope, lib, dom, b, 2, 27, (-7, 9))
        scope.check(lib, st)
        for h, shape, add in ((ha, a.shape, 100.0), (hb, b.shape, 200.0)):
            s, devs = begin_device(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL | scope.WRITE_ALL)])
            scope.check(lib, s)
            dev_view(devs[0], shape)[...] += add
            scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        before = bytes(dom.stats())
        after_a, keep_a = layout_of(scope, b, 27, (-7, 9))
        after_b, keep_b = layout_of(scope, a, 11, (-2, 3))
        cur_a, cur_b = c_uint64(), c_uint64()
        scope.check(lib, lib.fort_scope_identity_swap_v1(dom.ctx, ha, hb, byref(after_a), byref(after_b),
                                                         byref(cur_a), byref(cur_b)))
        assert (cur_a.value, cur_b.value) == (hb, ha)
        assert bytes(dom.stats()) == before  # nothing copied, allocated or waited for
        # Identity 1 now names b's allocation, identity 2 a's, with their device values.
        st, again1, _ = register(scope, lib, dom, b, 1, 27, (-7, 9))
        scope.check(lib, st)
        st, again2, _ = register(scope, lib, dom, a, 2, 11, (-2, 3))
        scope.check(lib, st)
        assert (again1, again2) == (hb, ha)
        assert list(a.ravel(order="F")) == [1, 2, 3, 4]  # device-current values not yet published
    assert list(a.ravel(order="F")) == [101, 102, 103, 104]
    assert list(b.ravel(order="F")) == [211, 212, 213, 214, 215, 216]


@pytest.mark.parametrize("invalid", ["host", "generation", "extent", "lower", "same_handle", "prepared"])
def test_a_bad_identity_swap_fails_the_context_without_publication(scope, cpu_runtime, invalid):
    lib = cpu_runtime
    a, b = np.array([1.0, 2.0]), np.array([3.0, 4.0])
    dom = new_domain(scope, lib)
    st, ha, _ = register(scope, lib, dom, a, 1, 2, (0,))
    scope.check(lib, st)
    st, hb, _ = register(scope, lib, dom, b, 2, 7, (0,))
    scope.check(lib, st)
    s, devs = begin_device(scope, lib, dom, [binding(scope, ha, flags=scope.READ_ALL | scope.WRITE_ALL)])
    scope.check(lib, s)
    dev_view(devs[0], (2,))[0] = 1000.0
    scope.check(lib, lib.fort_scope_device_end(dom.ctx, ha))
    after_a, keep_a = layout_of(scope, b, 7, (0,))
    after_b, keep_b = layout_of(scope, a, 2, (0,))
    if invalid == "host":
        after_a.host = a.ctypes.data
    elif invalid == "generation":
        after_a.generation = 2
    elif invalid == "extent":
        after_a, keep_a = layout_of(scope, b[:1], 7, (0,))
        after_a.host = b.ctypes.data
    elif invalid == "lower":
        after_a, keep_a = layout_of(scope, b, 7, (-1,))
    elif invalid == "same_handle":
        hb = ha
    elif invalid == "prepared":
        scope.check(lib, begin_device(scope, lib, dom, [binding(scope, hb, flags=scope.READ_ALL)])[0])
    cur_a, cur_b = c_uint64(111), c_uint64(222)
    assert lib.fort_scope_identity_swap_v1(dom.ctx, ha, hb, byref(after_a), byref(after_b), byref(cur_a),
                                           byref(cur_b)) == scope.EXECUTION
    assert (cur_a.value, cur_b.value) == (111, 222)
    assert lib.fort_scope_wait(dom.ctx) == scope.EXECUTION
    assert lib.fort_scope_close(dom.ctx) == scope.EXECUTION
    scope.check(lib, lib.fort_scope_abandon(dom.ctx))
    scope.check(lib, lib.fort_scope_domain_abandon_v1(dom.domain))
    dom.closed = True
    assert list(a) == [1.0, 2.0] and list(b) == [3.0, 4.0]


# ---------------------------------------------------------------------------------------------------------
# Dead values (discard, recycle, retire: compiler/tests/test_scoped_graph_replay.py:126-171), reached
# source reads (compiler/tests/test_source_read_runtime.py), scratch and completion counters and the
# trace attribution state.

def register_defined(scope, lib, dom, array, boxes, lower):
    ext = (c_size_t * array.ndim)(*array.shape)
    low = (ctypes.c_int64 * array.ndim)(*lower)
    layout = scope.Layout(array.ndim, scope.REAL64, 8, array.ctypes.data, ext, low, 0)
    secs = scope.Sections(boxes)
    h = c_uint64()
    scope.check(lib, lib.fort_scope_register_sections(dom.ctx, 77, 0, byref(layout), secs.array, secs.count,
                                                      byref(h)))
    return h.value


def admit(scope, lib, dom, h, boxes):
    secs = scope.Sections(boxes)
    return lib.fort_scope_admit_source_reads_v1(dom.ctx, h, secs.array, secs.count)


def host_read(scope, lib, dom, h, flags=0, reads=()):
    access, _keep = scope.make_access(flags, reads, (), ())
    return lib.fort_scope_host_begin(dom.ctx, h, byref(access))


def test_discard_and_retire_publish_no_dead_values(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL | scope.WRITE_ALL)])
        scope.check(lib, st)
        dev_view(devs[0], (8,))[0] = 1000.0
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        before = dom.stats()
        scope.check(lib, lib.fort_scope_discard_v1(dom.ctx, h))
        assert dom.stats().downloads == before.downloads and dom.stats().waits == before.waits
        assert host_read(scope, lib, dom, h, flags=scope.READ_ALL) == scope.UNINITIALIZED
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, flags=scope.WRITE_ALL | scope.OVERWRITE_ALL)])
        scope.check(lib, st)
        dev_view(devs[0], (8,))[:] = 2000.0 + np.arange(8)
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        assert dom.stats().allocations == 1
        scope.check(lib, lib.fort_scope_retire_v1(dom.ctx, h))
        assert dom.stats().downloads == 0 and dom.stats().allocated_bytes == 0
        assert host_read(scope, lib, dom, h, flags=scope.READ_ALL) == scope.STALE
    assert list(a) == list(range(8))


def test_recycle_without_queued_users_forgets_like_discard(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        scope.check(lib, lib.fort_scope_recycle_v1(dom.ctx, h))
        assert host_read(scope, lib, dom, h, flags=scope.READ_ALL) == scope.UNINITIALIZED
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, writes=[((0,), (8,))], overwrites=[((0,), (8,))])])
        scope.check(lib, st)
        assert lib.fort_scope_recycle_v1(dom.ctx, h) == scope.STATE  # a prepared access
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))


def test_retire_and_discard_refuse_a_prepared_access_and_retire_waits(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, flags=scope.READ_ALL)])
        scope.check(lib, st)
        assert lib.fort_scope_discard_v1(dom.ctx, h) == scope.STATE
        assert lib.fort_scope_retire_v1(dom.ctx, h) == scope.STATE
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        before = dom.stats().waits
        scope.check(lib, lib.fort_scope_observe_v1(dom.ctx, None, 0))
        assert dom.stats().waits == before
        scope.check(lib, lib.fort_scope_retire_v1(dom.ctx, h))
        assert dom.stats().waits == before + 1


def test_admitted_source_reads_keep_dirty_device_values_and_unread_cells(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(8.0)
    dom = new_domain(scope, lib)
    h = register_defined(scope, lib, dom, a, [((2,), (4,))], (-3,))
    st, devs = begin_device(scope, lib, dom, [binding(scope, h, reads=[((2,), (4,))], writes=[((2,), (4,))])])
    scope.check(lib, st)
    dev_view(devs[0], (8,))[2:4] = [102.0, 103.0]
    scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
    before = bytes(dom.stats())
    scope.check(lib, admit(scope, lib, dom, h, [((1,), (6,))]))
    scope.check(lib, admit(scope, lib, dom, h, [((1,), (6,))]))
    assert bytes(dom.stats()) == before  # no transfer, no wait
    # A repeated admission must not mark the stale host copy current over the device writes.
    st, devs = begin_device(scope, lib, dom, [binding(scope, h, reads=[((1,), (6,))])])
    scope.check(lib, st)
    assert list(dev_view(devs[0], (8,))[1:6]) == [1, 102, 103, 4, 5]
    scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
    assert dom.stats().upload_bytes == 5 * 8
    assert host_read(scope, lib, dom, h, reads=[((0,), (1,))]) == scope.UNINITIALIZED
    dom.close()
    assert list(a) == [0, 1, 102, 103, 4, 5, 6, 7]
    assert admit(scope, lib, dom, h, [((0,), (1,))]) == scope.STALE


def test_source_read_admission_bounds_prepared_access_and_region_limit(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        assert admit(scope, lib, dom, h, [((0,), (9,))]) == scope.ARGUMENT
        scope.check(lib, admit(scope, lib, dom, h, [((0,), (1,))]))
        access, keep = scope.make_access(0, [((0,), (1,))], (), ())
        scope.check(lib, lib.fort_scope_host_begin(dom.ctx, h, byref(access)))
        assert admit(scope, lib, dom, h, [((1,), (2,))]) == scope.STATE
        assert b"prepared" in lib.fort_scope_error()
        scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
        assert host_read(scope, lib, dom, h, reads=[((1,), (2,))]) == scope.UNINITIALIZED
    limit = 128  # rectangles per region
    b = np.arange(3.0 * limit + 4)
    dom = new_domain(scope, lib)
    h = register_defined(scope, lib, dom, b, [((3 * i,), (3 * i + 1,)) for i in range(limit)], (1,))
    before = bytes(dom.stats())
    assert admit(scope, lib, dom, h, [((3 * limit,), (3 * limit + 1,))]) == scope.BOUNDARY
    assert bytes(dom.stats()) == before
    assert host_read(scope, lib, dom, h, reads=[((3 * limit,), (3 * limit + 1,))]) == scope.UNINITIALIZED
    dom.close()


def test_scratch_and_completion_counters(scope, cpu_runtime):
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        lease = scope.ScratchLease()
        for size in (1024, 512, 4096):
            scope.check(lib, lib.fort_scope_scratch_acquire_v1(dom.ctx, size, byref(lease)))
            s = scope.ScratchStats()
            scope.check(lib, lib.fort_scope_scratch_stats_get_v1(dom.ctx, byref(s)))
            assert s.active == 1 and s.active_bytes == size
            scope.check(lib, lib.fort_scope_scratch_release_v1(dom.ctx, lease.token))
        scope.check(lib, lib.fort_scope_scratch_stats_get_v1(dom.ctx, byref(s)))
        assert (s.version, s.active, s.acquisitions, s.allocations, s.reuses, s.releases, s.grows) == \
            (1, 0, 3, 2, 1, 3, 1)
        assert s.capacity_bytes == 4096 and s.peak_scratch_bytes == 4096 and s.active_bytes == 0
        h = dom.bind(a)
        cs = scope.CompletionStats()
        scope.check(lib, lib.fort_scope_completion_stats_get_v1(dom.ctx, byref(cs)))
        independent = cs.independent_host_tasks
        assert host_read(scope, lib, dom, h, flags=scope.READ_ALL) == scope.OK  # current on the host
        scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
        scope.check(lib, lib.fort_scope_completion_stats_get_v1(dom.ctx, byref(cs)))
        assert cs.version == 1 and cs.independent_host_tasks == independent + 1
        assert cs.event_records == cs.event_waits == 0


def test_trace_attribution_is_saved_and_restored_only_while_tracing(scope, cpu_runtime, monkeypatch):
    lib = cpu_runtime
    with new_domain(scope, lib) as dom:
        monkeypatch.delenv("FORT_RUNTIME_TRACE", raising=False)
        prev = scope.TraceState()
        prev.version = 77
        lib.fort_scope_trace_set_v1(dom.ctx, 63, b"o", b"p", b"s", b"op", b"impl", b"b", byref(prev))
        assert prev.version == 77  # tracing off: returns before looking at the context
        monkeypatch.setenv("FORT_RUNTIME_TRACE", "1")
        lib.fort_scope_trace_set_v1(dom.ctx, 1 | 2, b"owner1", b"proc1", None, None, None, None, byref(prev))
        assert prev.owner == b"" and prev.procedure == b""
        saved = scope.TraceState()
        lib.fort_scope_trace_set_v1(dom.ctx, 2, None, b"proc2", None, None, None, None, byref(saved))
        assert saved.owner == b"owner1" and saved.procedure == b"proc1"
        now = scope.TraceState()
        lib.fort_scope_trace_set_v1(dom.ctx, 0, None, None, None, None, None, None, byref(now))  # mask 0 saves
        assert now.owner == b"owner1" and now.procedure == b"proc2"
        lib.fort_scope_trace_restore_v1(dom.ctx, byref(saved))
        lib.fort_scope_trace_set_v1(dom.ctx, 0, None, None, None, None, None, None, byref(now))
        assert now.procedure == b"proc1"
        too_long = b"x" * 65  # no NUL within 65 bytes: unknown
        lib.fort_scope_trace_set_v1(dom.ctx, 1, too_long, None, None, None, None, None, byref(now))
        lib.fort_scope_trace_set_v1(dom.ctx, 0, None, None, None, None, None, None, byref(now))
        assert now.owner == b""


class Kernel(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("function", c_void_p),
                ("grid", ctypes.c_uint32 * 3), ("block", ctypes.c_uint32 * 3), ("shared_bytes", c_size_t),
                ("arguments", c_void_p), ("argument_count", c_size_t)]


class GraphTask(ctypes.Structure):
    _fields_ = [("kernel", Kernel), ("bindings", c_void_p), ("binding_count", c_size_t)]


def group_task(scope, uses, keep):
    arr = (scope.PlanBinding * len(uses))(*[scope.PlanBinding(h, acc) for h, acc in uses])
    keep.append(arr)
    k = Kernel(1, 0, c_void_p(0x1000), (ctypes.c_uint32 * 3)(1, 1, 1), (ctypes.c_uint32 * 3)(64, 1, 1), 0, None, 0)
    return GraphTask(k, ctypes.cast(arr, c_void_p), len(uses))


def test_group_execution_simulates_effects_in_order(scope, cpu_runtime):
    # compiler/tests/test_scoped_graph_replay.py:172-194: a producer may define what a later task reads;
    # the CPU backend validates the whole group, then refuses to run it (RESOURCE) and commits nothing.
    lib = cpu_runtime
    arrays = [np.arange(8.0) for _ in range(3)]
    with new_domain(scope, lib) as dom:
        a = dom.bind(arrays[0])
        b = dom.bind(arrays[1], initialized=False)
        out = dom.bind(arrays[2], initialized=False)
        keep = []
        ra, _k1 = scope.make_access(scope.READ_ALL)
        rb, _k2 = scope.make_access(scope.READ_ALL)
        wb, _k3 = scope.make_access(scope.WRITE_ALL | scope.OVERWRITE_ALL)
        wo, _k4 = scope.make_access(scope.WRITE_ALL | scope.OVERWRITE_ALL)
        produce = group_task(scope, [(a, ra), (b, wb)], keep)
        consume = group_task(scope, [(b, rb), (out, wo)], keep)
        before = bytes(dom.stats())
        ordered = (GraphTask * 2)(produce, consume)
        assert lib.fort_scope_graph_execute_v1(dom.ctx, 8, ordered, 2) == scope.RESOURCE
        assert bytes(dom.stats()) == before
        assert host_read(scope, lib, dom, b, flags=scope.READ_ALL) == scope.UNINITIALIZED
        assert host_read(scope, lib, dom, out, flags=scope.READ_ALL) == scope.UNINITIALIZED
        reverse = (GraphTask * 2)(consume, produce)
        assert lib.fort_scope_graph_execute_v1(dom.ctx, 8, reverse, 2) == scope.UNINITIALIZED
        stale = group_task(scope, [(2**63, ra)], keep)
        assert lib.fort_scope_graph_execute_v1(dom.ctx, 9, (GraphTask * 2)(produce, stale), 2) == scope.STALE
        assert bytes(dom.stats()) == before
        assert lib.fort_scope_graph_execute_v1(dom.ctx, 9, None, 0) == scope.OK  # an empty group
        assert lib.fort_scope_graph_submit_v1(dom.ctx, 1, None, 0) == scope.OK
        assert lib.fort_scope_graph_submit_v1(dom.ctx, 1, ctypes.byref(produce.kernel), 1) == scope.RESOURCE


def test_a_definition_kill_keeps_the_physical_values_for_publication(scope, cpu_runtime):
    # compiler/tests/test_scoped_runtime.py:509-531: forgetting a definition refuses later reads but keeps
    # the device's values as the current physical copy, so closing still publishes them.
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        every = scope.WRITE_ALL | scope.OVERWRITE_ALL
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, flags=every)])
        scope.check(lib, st)
        dev_view(devs[0], (8,))[:] = 1000.0 + np.arange(8)
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        scope.check(lib, lib.fort_scope_forget_definition(dom.ctx, h))
        assert dom.stats().downloads == 0
        assert host_read(scope, lib, dom, h, flags=scope.READ_ALL) == scope.UNINITIALIZED
        middle = ((2,), (6,))
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, writes=[middle], overwrites=[middle])])
        scope.check(lib, st)
        dev_view(devs[0], (8,))[2:6] = 2000.0 + np.arange(2, 6)
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        assert dom.stats().allocations == 1
    assert list(a) == [1000, 1001, 2002, 2003, 2004, 2005, 1006, 1007]


def test_a_definition_kill_cannot_lose_an_executed_device_prefix(scope, cpu_runtime):
    # compiler/tests/test_scoped_runtime.py:533-560
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        written = ((1,), (7,))
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, writes=[written], overwrites=[written])])
        scope.check(lib, st)
        dev_view(devs[0], (8,))[1:7] = 3000.0 + np.arange(1, 7)
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        before = dom.stats()
        scope.check(lib, lib.fort_scope_forget_definition(dom.ctx, h))
        after = dom.stats()
        assert (after.download_bytes, after.waits) == (before.download_bytes, before.waits)
        assert host_read(scope, lib, dom, h, reads=[((3,), (4,))]) == scope.UNINITIALIZED
    assert list(a) == [0, 3001, 3002, 3003, 3004, 3005, 3006, 7]


def test_a_partial_write_needs_the_values_it_preserves(scope, cpu_runtime):
    # The cells a write may leave as they were must be defined (compiler/runtime/scoped_runtime.cu:925-928),
    # on the device and on the host alike; what it overwrites needs nothing.
    lib = cpu_runtime
    a = np.zeros(8)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        partial = binding(scope, h, writes=[((0,), (8,))], overwrites=[((2,), (6,))])
        st, _ = begin_device(scope, lib, dom, [partial])
        assert st == scope.UNINITIALIZED
        access, _keep = scope.make_access(0, (), [((0,), (8,))], [((2,), (6,))])
        assert lib.fort_scope_host_begin(dom.ctx, h, byref(access)) == scope.UNINITIALIZED
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, writes=[((2,), (6,))], overwrites=[((2,), (6,))])])
        scope.check(lib, st)
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, writes=[((2,), (6,))])])
        scope.check(lib, st)  # now defined: a partial write inside it is accepted
        scope.check(lib, lib.fort_scope_cancel_access(dom.ctx, h))


def test_admitted_reads_after_a_kill_keep_the_newer_device_values(scope, cpu_runtime):
    # compiler/runtime/scoped_regions.hpp:205-216: a value the device physically holds keeps that copy;
    # only values no copy tracks take the host's.
    lib = cpu_runtime
    a = np.arange(8.0)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        part = ((0,), (4,))
        st, devs = begin_device(scope, lib, dom, [binding(scope, h, writes=[part], overwrites=[part])])
        scope.check(lib, st)
        dev_view(devs[0], (8,))[0:4] = 50.0
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
        scope.check(lib, lib.fort_scope_forget_definition(dom.ctx, h))
        scope.check(lib, admit(scope, lib, dom, h, [((0,), (8,))]))
        scope.check(lib, host_read(scope, lib, dom, h, flags=scope.READ_ALL))
        assert list(a) == [50, 50, 50, 50, 4, 5, 6, 7]
        assert dom.stats().download_bytes == 4 * 8
        scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))


def test_fragmentation_reconciles_current_values(scope, cpu_runtime):
    # compiler/tests/test_scoped_runtime.py:293-315: alternating single-cell host and device writes need
    # more than 128 rectangles; the runtime reconciles instead of failing, and closing publishes every value.
    lib = cpu_runtime
    size = 4 * 128
    a = np.arange(float(size))
    expected = a.copy()
    with new_domain(scope, lib) as dom:
        h = dom.bind(a)
        for i in range(2 * 128 + 16):
            box = ((i,), (i + 1,))
            if i % 2:
                access, _keep = scope.make_access(0, (), [box], [box])
                scope.check(lib, lib.fort_scope_host_begin(dom.ctx, h, byref(access)))
                a[i] = -100.0 - i
                scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
                expected[i] = -100.0 - i
            else:
                st, devs = begin_device(scope, lib, dom, [binding(scope, h, writes=[box], overwrites=[box])])
                scope.check(lib, st)
                dev_view(devs[0], (size,))[i] = 1000.0 + i
                scope.check(lib, lib.fort_scope_device_end(dom.ctx, h))
                expected[i] = 1000.0 + i
        assert dom.stats().reconciliations > 0
    assert np.array_equal(a, expected)


def test_a_definition_beyond_the_rectangle_limit_is_a_boundary_before_execution(scope, cpu_runtime):
    # compiler/tests/test_scoped_runtime.py:341-357: 128 disjoint defined cells, and the 129th host
    # overwrite is refused with BOUNDARY at preparation: nothing committed, no access left open.
    lib = cpu_runtime
    size = 2 * 128 + 16
    a = np.zeros(size)
    with new_domain(scope, lib) as dom:
        h = dom.bind(a, initialized=False)
        for i in range(128):
            box = ((2 * i,), (2 * i + 1,))
            access, _keep = scope.make_access(0, (), [box], [box])
            scope.check(lib, lib.fort_scope_host_begin(dom.ctx, h, byref(access)))
            a[2 * i] = i + 10
            scope.check(lib, lib.fort_scope_host_end(dom.ctx, h))
        box = ((256,), (257,))
        access, _keep = scope.make_access(0, (), [box], [box])
        assert lib.fort_scope_host_begin(dom.ctx, h, byref(access)) == scope.BOUNDARY
        st, _ = begin_device(scope, lib, dom, [binding(scope, h, writes=[box], overwrites=[box])])
        assert st == scope.BOUNDARY
    assert list(a[0:256:2]) == [i + 10 for i in range(128)] and a[256] == 0


def test_a_read_upload_past_the_rectangle_limit_reconciles_before_any_binding_changes(scope, cpu_runtime):
    # compiler/runtime/scoped_runtime.cu:899-954: the acquisition of every binding, transfers included, is
    # simulated before anything moves, and a read transfer that would split the device currency past 128
    # rectangles reconciles (the device takes every host value) instead of failing half-way with the first
    # binding already prepared.
    lib = cpu_runtime
    a = np.arange(16.0)
    b = np.arange(4.0 * 200) + 0.5
    with new_domain(scope, lib) as dom:
        ha, hb = dom.bind(a), dom.bind(b)
        for i in range(127):
            box = ((4 * i,), (4 * i + 1,))
            st, _ = begin_device(scope, lib, dom, [binding(scope, hb, reads=[box])])
            scope.check(lib, st)
            scope.check(lib, lib.fort_scope_device_end(dom.ctx, hb))
        before = dom.stats().reconciliations
        st, devs = begin_device(scope, lib, dom, [binding(scope, ha, flags=scope.WRITE_ALL),
                                                  binding(scope, hb, reads=[((2,), (3,)), ((6,), (7,))])])
        scope.check(lib, st)
        assert dom.stats().reconciliations == before + 1
        # The device copy of b holds every host value now, and a is writable.
        assert np.array_equal(dev_view(devs[1], b.shape), b)
        dev_view(devs[0], a.shape)[:] = -a
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, ha))
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, hb))
    assert np.array_equal(a, -np.arange(16.0)) and np.array_equal(b, np.arange(4.0 * 200) + 0.5)


def test_a_refused_later_binding_leaves_the_earlier_buffer_unprepared(scope, cpu_runtime):
    # A two-binding task whose second binding cannot be committed (a 129th disjoint definition) is refused
    # with BOUNDARY before the first binding changes: the first buffer has no prepared access afterwards and
    # its host values stand.
    lib = cpu_runtime
    a = np.arange(8.0)
    size = 2 * 128 + 16
    b = np.zeros(size)
    with new_domain(scope, lib) as dom:
        ha, hb = dom.bind(a), dom.bind(b, initialized=False)
        for i in range(128):
            box = ((2 * i,), (2 * i + 1,))
            access, _keep = scope.make_access(0, (), [box], [box])
            scope.check(lib, lib.fort_scope_host_begin(dom.ctx, hb, byref(access)))
            b[2 * i] = i + 10
            scope.check(lib, lib.fort_scope_host_end(dom.ctx, hb))
        box = ((256,), (257,))
        st, _ = begin_device(scope, lib, dom, [binding(scope, ha, flags=scope.WRITE_ALL),
                                               binding(scope, hb, writes=[box], overwrites=[box])])
        assert st == scope.BOUNDARY
        st, devs = begin_device(scope, lib, dom, [binding(scope, ha, flags=scope.READ_ALL)])
        scope.check(lib, st)
        assert np.array_equal(dev_view(devs[0], a.shape), np.arange(8.0))
        scope.check(lib, lib.fort_scope_device_end(dom.ctx, ha))
    assert np.array_equal(a, np.arange(8.0)) and list(b[0:256:2]) == [i + 10 for i in range(128)]
