"""ctypes bindings of the fort_scope C ABI (include/fort_scope.h) and a small host-side mirror of the
reference's Fortran ``resident_domain`` type (compiler/runtime/resident_domain.f90:162-976): a domain
with one context, arrays bound by identity, device-resident entries, observation and close.

The same code drives the HIP library (libftc_hip.so) and the CPU reference backend (libftc_cpu.so); the
latter holds no kernels and refuses device tasks with RESOURCE, as the reference's CPU test backend does.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int64, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

OK, ARGUMENT, STALE, ALIAS, RESOURCE, BOUNDARY, EXECUTION, UNINITIALIZED, STATE, DECLINED = range(10)
STATUS_NAMES = ("OK", "ARGUMENT", "STALE", "ALIAS", "RESOURCE", "BOUNDARY", "EXECUTION", "UNINITIALIZED",
                "STATE", "DECLINED")
BYTES, REAL32, REAL64, INTEGER32, LOGICAL = range(5)
READ_ALL, WRITE_ALL, OVERWRITE_ALL = 1, 2, 4
NATIVE, GPU = 0, 1


class Section(ctypes.Structure):
    _fields_ = [("lower", POINTER(c_size_t)), ("upper", POINTER(c_size_t))]


class Access(ctypes.Structure):
    _fields_ = [("flags", c_uint32), ("read_count", c_size_t), ("reads", POINTER(Section)),
                ("write_count", c_size_t), ("writes", POINTER(Section)),
                ("overwrite_count", c_size_t), ("overwrites", POINTER(Section))]


class Layout(ctypes.Structure):
    _fields_ = [("rank", c_uint32), ("type", c_uint32), ("element_bytes", c_size_t), ("host", c_void_p),
                ("extents", POINTER(c_size_t)), ("lower_bounds", POINTER(c_int64)), ("generation", c_uint64)]


class Stats(ctypes.Structure):
    _fields_ = [(n, c_uint64) for n in ("uploads", "downloads", "upload_bytes", "download_bytes", "allocations",
                                        "allocated_bytes", "peak_device_bytes", "launches", "waits",
                                        "reconciliations")]


class PlanBinding(ctypes.Structure):
    _fields_ = [("buffer", c_uint64), ("access", Access)]


class ResidencyStats(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("reserved", c_uint32)] + [
        (n, c_uint64) for n in ("reserve_batches", "requested_resources", "reuse_hits", "evictions",
                                "evicted_bytes", "dirty_evictions", "dirty_download_bytes", "pressure_flushes")]


class GraphStats(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("reserved", c_uint32)] + [
        (n, c_uint64) for n in ("creations", "instantiations", "rebuilds", "submissions", "replays",
                                "parameter_updates", "logical_kernels", "live_graphs", "destructions")]


class QueueStats(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("queue_enabled", c_uint32), ("graphs_enabled", c_uint32),
                ("reserved", c_uint32)] + [
        (n, c_uint64) for n in ("accepted_tasks", "submissions", "conflict_flushes", "explicit_flushes",
                                "pending_tasks", "peak_pending_tasks", "copied_argument_bytes",
                                "direct_submissions", "direct_kernels", "graph_resource_fallbacks")]


class ScratchLease(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("reserved", c_uint32), ("token", c_uint64), ("device", c_void_p),
                ("bytes", c_size_t), ("capacity", c_size_t)]


class ScratchStats(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("active", c_uint32)] + [
        (n, c_uint64) for n in ("acquisitions", "allocations", "reuses", "releases", "grows", "capacity_bytes",
                                "active_bytes", "peak_scratch_bytes", "peak_total_device_bytes")]


class CompletionStats(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("reserved", c_uint32)] + [
        (n, c_uint64) for n in ("event_records", "event_queries", "event_waits", "stream_waits", "created_events",
                                "reused_events", "independent_host_tasks")]


class TraceState(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("reserved", c_uint32)] + [
        (n, ctypes.c_char * 65) for n in ("owner", "procedure", "segment", "operation", "implementation", "boundary")]


class ExternalBinding(ctypes.Structure):
    _fields_ = [("buffer", c_uint64), ("type", c_uint32), ("rank", c_uint32), ("element_bytes", c_size_t),
                ("access", Access)]


class ExternalView(ctypes.Structure):
    _fields_ = [("buffer", c_uint64), ("generation", c_uint64), ("type", c_uint32), ("rank", c_uint32),
                ("element_bytes", c_size_t), ("bytes", c_size_t), ("extents", POINTER(c_size_t)),
                ("strides", POINTER(c_size_t)), ("lower_bounds", POINTER(c_int64)), ("device", c_void_p)]


class ExternalCall(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("reserved", c_uint32), ("bindings", POINTER(ExternalBinding)),
                ("binding_count", c_size_t), ("workspace_lease", c_uint64), ("workspace_bytes", c_size_t)]


class ExternalReport(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("status", ctypes.c_int32), ("started", c_uint32), ("committed", c_uint32)]


EXTERNAL_WORKER = ctypes.CFUNCTYPE(c_int, c_void_p, POINTER(ExternalView), c_size_t, c_void_p, POINTER(c_int))

# name -> (restype, argtypes) of every runtime function the ABI header declares.
U64P = POINTER(c_uint64)
RUNTIME_SIGNATURES = {
    "fort_scope_abi_version": (c_uint32, []),
    "fort_scope_error": (c_char_p, []),
    "fort_scope_create": (c_int, [c_int, U64P]),
    "fort_scope_domain_create_v1": (c_int, [c_int, U64P]),
    "fort_scope_create_in_domain_v1": (c_int, [c_uint64, U64P]),
    "fort_scope_domain_get_v1": (c_int, [c_uint64, U64P]),
    "fort_scope_domain_device_get_v1": (c_int, [c_uint64, POINTER(c_int)]),
    "fort_scope_domain_stats_get_v1": (c_int, [c_uint64, POINTER(Stats)]),
    "fort_scope_domain_close_v1": (c_int, [c_uint64]),
    "fort_scope_domain_abandon_v1": (c_int, [c_uint64]),
    "fort_scope_domain_workspace_reserve_v1": (c_int, [c_uint64, c_size_t, U64P]),
    "fort_scope_domain_workspace_resize_v1": (c_int, [c_uint64, c_uint64, c_size_t]),
    "fort_scope_domain_workspace_release_v1": (c_int, [c_uint64, c_uint64]),
    "fort_scope_set_device_budget": (c_int, [c_uint64, c_size_t]),
    "fort_scope_device_budget_default_v1": (c_int, [c_int, POINTER(c_size_t)]),
    "fort_scope_domain_budget_get_v1": (c_int, [c_uint64, POINTER(c_size_t), POINTER(c_size_t)]),
    "fort_scope_scratch_acquire_v1": (c_int, [c_uint64, c_size_t, POINTER(ScratchLease)]),
    "fort_scope_scratch_release_v1": (c_int, [c_uint64, c_uint64]),
    "fort_scope_register": (c_int, [c_uint64, c_uint64, c_uint64, POINTER(Layout), c_int, U64P]),
    "fort_scope_register_sections": (c_int, [c_uint64, c_uint64, c_uint64, POINTER(Layout), POINTER(Section),
                                             c_size_t, U64P]),
    "fort_scope_forget_definition": (c_int, [c_uint64, c_uint64]),
    "fort_scope_layout_get": (c_int, [c_uint64, c_uint64, POINTER(Layout)]),
    "fort_scope_access_begin_v1": (c_int, [c_uint64, POINTER(PlanBinding), c_size_t, c_int, POINTER(c_void_p)]),
    "fort_scope_original_host_begin_v1": (c_int, [c_uint64, POINTER(PlanBinding), c_size_t]),
    "fort_scope_graph_submit_v1": (c_int, [c_uint64, c_uint64, c_void_p, c_size_t]),
    "fort_scope_graph_execute_v1": (c_int, [c_uint64, c_uint64, c_void_p, c_size_t]),
    "fort_scope_recycle_v1": (c_int, [c_uint64, c_uint64]),
    "fort_scope_retire_v1": (c_int, [c_uint64, c_uint64]),
    "fort_scope_admit_source_reads_v1": (c_int, [c_uint64, c_uint64, POINTER(Section), c_size_t]),
    "fort_scope_scratch_stats_get_v1": (c_int, [c_uint64, POINTER(ScratchStats)]),
    "fort_scope_completion_stats_get_v1": (c_int, [c_uint64, POINTER(CompletionStats)]),
    "fort_scope_trace_set_v1": (None, [c_uint64, c_uint32] + [ctypes.c_char_p] * 6 + [POINTER(TraceState)]),
    "fort_scope_trace_restore_v1": (None, [c_uint64, POINTER(TraceState)]),
    "fort_scope_identity_swap_v1": (c_int, [c_uint64, c_uint64, c_uint64, POINTER(Layout), POINTER(Layout), U64P,
                                            U64P]),
    "fort_scope_external_execute_v1": (c_int, [c_uint64, POINTER(ExternalCall), EXTERNAL_WORKER, c_void_p,
                                               POINTER(ExternalReport)]),
    "fort_scope_observe_v1": (c_int, [c_uint64, U64P, c_size_t]),
    "fort_scope_observe_all_v1": (c_int, [c_uint64]),
    "fort_scope_discard_v1": (c_int, [c_uint64, c_uint64]),
    "fort_scope_host_begin": (c_int, [c_uint64, c_uint64, POINTER(Access)]),
    "fort_scope_host_end": (c_int, [c_uint64, c_uint64]),
    "fort_scope_device_begin": (c_int, [c_uint64, c_uint64, POINTER(Access), POINTER(c_void_p)]),
    "fort_scope_device_end": (c_int, [c_uint64, c_uint64]),
    "fort_scope_device_reserve_v1": (c_int, [c_uint64, c_uint64, POINTER(c_void_p)]),
    "fort_scope_device_reserve_many_v1": (c_int, [c_uint64, U64P, c_size_t, POINTER(c_void_p)]),
    "fort_scope_buffer_valid_v1": (c_int, [c_uint64, c_uint64, POINTER(c_int)]),
    "fort_scope_residency_stats_get_v1": (c_int, [c_uint64, POINTER(ResidencyStats)]),
    "fort_scope_cancel_access": (c_int, [c_uint64, c_uint64]),
    "fort_scope_gpu_enter": (c_int, [c_uint64, POINTER(c_int), POINTER(c_void_p)]),
    "fort_scope_gpu_leave": (c_int, [c_uint64, c_int]),
    "fort_scope_note_launch": (c_int, [c_uint64]),
    "fort_scope_graph_enqueue_v1": (c_int, [c_uint64, c_uint64, c_void_p, POINTER(c_size_t)]),
    "fort_scope_graph_enqueue_members_v1": (c_int, [c_uint64, c_uint64, c_void_p, POINTER(c_size_t), c_void_p,
                                                    c_size_t]),
    "fort_scope_graph_flush_v1": (c_int, [c_uint64]),
    "fort_scope_graph_queue_cancel_v1": (c_int, [c_uint64]),
    "fort_scope_graph_queue_stats_get_v1": (c_int, [c_uint64, POINTER(QueueStats)]),
    "fort_scope_graph_stats_get_v1": (c_int, [c_uint64, POINTER(GraphStats)]),
    "fort_scope_execution_error": (c_int, [c_uint64, c_char_p]),
    "fort_scope_report_error": (c_int, [c_int, c_char_p]),
    "fort_scope_wait": (c_int, [c_uint64]),
    "fort_scope_stats_get": (c_int, [c_uint64, POINTER(Stats)]),
    "fort_scope_unregister": (c_int, [c_uint64, c_uint64]),
    "fort_scope_close": (c_int, [c_uint64]),
    "fort_scope_abandon": (c_int, [c_uint64]),
}


class ScopeError(RuntimeError):
    def __init__(self, status: int, message: str):
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"{name}: {message}")
        self.status = status


def bind_runtime(lib: ctypes.CDLL) -> ctypes.CDLL:
    for name, (res, args) in RUNTIME_SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def check(lib: ctypes.CDLL, status: int) -> int:
    if status not in (OK, DECLINED):
        raise ScopeError(status, (lib.fort_scope_error() or b"").decode())
    return status


def _sizes(values) -> ctypes.Array:
    return (c_size_t * len(values))(*values)


class Sections:
    """Keeps the ctypes storage of a list of zero-based half-open boxes alive."""

    def __init__(self, boxes):
        self.bounds = [(_sizes(lo), _sizes(hi)) for lo, hi in boxes]
        self.array = (Section * max(1, len(boxes)))(*[Section(lo, hi) for lo, hi in self.bounds])
        self.count = len(boxes)


def make_access(flags=0, reads=(), writes=(), overwrites=()):
    """An Access whose section storage lives as long as the returned tuple."""
    r, w, o = Sections(list(reads)), Sections(list(writes)), Sections(list(overwrites))
    access = Access(flags, r.count, r.array if r.count else None, w.count, w.array if w.count else None,
                    o.count, o.array if o.count else None)
    return access, (r, w, o)


_TYPES = {np.dtype(np.float64): REAL64, np.dtype(np.float32): REAL32, np.dtype(np.int32): INTEGER32}


class Domain:
    """A residency domain with one borrowing context: the host-side counterpart of the reference's
    Fortran ``resident_domain`` (begin_epoch, bind_*, observe, end_epoch, close)."""

    def __init__(self, lib: ctypes.CDLL, device: int = 0, budget: int | None = None):
        self.lib = lib
        self.domain = c_uint64()
        check(lib, lib.fort_scope_domain_create_v1(device, ctypes.byref(self.domain)))
        self.context = c_uint64()
        check(lib, lib.fort_scope_create_in_domain_v1(self.domain, ctypes.byref(self.context)))
        if budget is not None:
            check(lib, lib.fort_scope_set_device_budget(self.context, budget))
        self.arrays: dict[int, tuple[np.ndarray, object]] = {}
        self._next_identity = 1
        self.closed = False

    @property
    def ctx(self) -> int:
        return self.context.value

    def bind(self, array: np.ndarray, initialized: bool = True, identity: int | None = None, lower=None) -> int:
        """Registers a Fortran-ordered host array (borrowed until close) and returns its handle. `lower`
        gives the Fortran lower bounds (default 1 in every dimension)."""
        if not array.flags.f_contiguous:
            raise ValueError("arrays must be Fortran-contiguous")
        kind = _TYPES.get(array.dtype, BYTES)
        ext = _sizes(array.shape)
        lower = (c_int64 * array.ndim)(*(lower if lower is not None else [1] * array.ndim))
        layout = Layout(array.ndim, kind, array.itemsize, array.ctypes.data, ext, lower, 0)
        handle = c_uint64()
        if identity is None:
            identity = self._next_identity
            self._next_identity += 1
        check(self.lib, self.lib.fort_scope_register(self.context, identity, 0, ctypes.byref(layout),
                                                     1 if initialized else 0, ctypes.byref(handle)))
        self.arrays[handle.value] = (array, (ext, lower))
        return handle.value

    def observe(self, *handles: int) -> None:
        arr = (c_uint64 * len(handles))(*handles)
        check(self.lib, self.lib.fort_scope_observe_v1(self.context, arr, len(handles)))

    def flush(self) -> None:
        check(self.lib, self.lib.fort_scope_graph_flush_v1(self.context))

    def wait(self) -> None:
        check(self.lib, self.lib.fort_scope_wait(self.context))

    def stream(self) -> int:
        prev, stream = c_int(), c_void_p()
        check(self.lib, self.lib.fort_scope_gpu_enter(self.context, ctypes.byref(prev), ctypes.byref(stream)))
        check(self.lib, self.lib.fort_scope_gpu_leave(self.context, prev.value))
        return stream.value or 0

    def stats(self) -> Stats:
        s = Stats()
        check(self.lib, self.lib.fort_scope_stats_get(self.context, ctypes.byref(s)))
        return s

    def graph_stats(self) -> GraphStats:
        s = GraphStats()
        check(self.lib, self.lib.fort_scope_graph_stats_get_v1(self.context, ctypes.byref(s)))
        return s

    def queue_stats(self) -> QueueStats:
        s = QueueStats()
        check(self.lib, self.lib.fort_scope_graph_queue_stats_get_v1(self.context, ctypes.byref(s)))
        return s

    def close(self) -> None:
        """Closes the context, then the domain: every defined value is published to its host array."""
        if self.closed:
            return
        self.closed = True
        check(self.lib, self.lib.fort_scope_close(self.context))
        check(self.lib, self.lib.fort_scope_domain_close_v1(self.domain))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if exc[0] is None:
            self.close()
        else:
            try:
                self.close()
            except ScopeError:
                pass
