"""The coherence rules of the runtime against a brute-force model on small arrays.

Each cell of a buffer carries three flags (defined, current on the host, current on the device) and a
value per copy. Random sequences of accesses drive the library and the model side by side: device and
host accesses with read, may-write and must-overwrite boxes, original host work, observation, definition
kills and discards. After every operation the statuses, the exact transfer byte counts and the host
array must agree with the model, and a device access must see the model's device values on what it reads.

The model restates compiler/runtime/scoped_runtime.cu:811-818 (effects), :899-954 (what an access
copies, and the definitions it needs) and compiler/runtime/scoped_regions.hpp:189-236 (physical values,
prepare, original host work, admitted reads); publication is :1410-1413. It runs on the CPU backend,
where "device" memory is host memory the test can read and write through the pointers the runtime
returns, and on the HIP backend, where the test reads and writes device memory by synchronous copies.
"""
import ctypes
from ctypes import byref, c_void_p

import numpy as np
import pytest

SHAPE = (6, 5)


def random_box(rng, allow_empty=True):
    lo, hi = [], []
    for n in SHAPE:
        a, b = sorted(rng.integers(0, n + 1, 2))
        if not allow_empty and a == b:
            a, b = (a, a + 1) if a < n else (a - 1, a)
        lo.append(int(a))
        hi.append(int(b))
    return tuple(lo), tuple(hi)


def mask(box):
    m = np.zeros(SHAPE, dtype=bool)
    (l0, l1), (h0, h1) = box
    m[l0:h0, l1:h1] = True
    return m


class Model:
    def __init__(self, host, initialized):
        self.init = np.full(SHAPE, initialized)
        self.hc = self.init.copy()
        self.dc = np.zeros(SHAPE, dtype=bool)
        self.host = host.copy()
        self.dev = np.full(SHAPE, np.nan)
        self.up = self.down = 0

    def physical(self):
        return self.init | self.hc | self.dc

    def plan(self, reads, writes, overwrites, device, strict=True):
        """The cells the access copies, or the status that refuses it."""
        if (overwrites & ~writes).any():
            return "ARGUMENT"
        wanted = reads | (writes & ~overwrites)
        if strict and (wanted & ~self.init).any():
            return "UNINITIALIZED"
        known = self.init if strict else self.physical()
        have = self.dc if device else self.hc
        return wanted & known & ~have

    def copy(self, cells, device):
        if device:
            self.dev[cells] = self.host[cells]
            self.dc |= cells
            self.up += int(cells.sum()) * 8
        else:
            self.host[cells] = self.dev[cells]
            self.hc |= cells
            self.down += int(cells.sum()) * 8

    def commit(self, writes, overwrites, device, original=False):
        self.init |= overwrites
        if original:
            self.hc |= overwrites
            self.dc &= ~writes
            return
        self.init |= writes
        if device:
            self.dc |= writes
            self.hc &= ~writes
        else:
            self.hc |= writes
            self.dc &= ~writes


class HostMemory:
    """Device memory of the CPU backend: host memory, read and written in place."""

    @staticmethod
    def read(ptr):
        n = int(np.prod(SHAPE))
        return np.frombuffer((ctypes.c_double * n).from_address(ptr), dtype=np.float64).reshape(SHAPE, order="F")

    @staticmethod
    def write(ptr, values):
        HostMemory.read(ptr)[...] = values


class DeviceMemory:
    """Device memory of the HIP backend, read and written by synchronous copies after the device is idle
    (the runtime's own copies run on its stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMemcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, ctypes.c_int]

    def read(self, ptr):
        assert self.hip.hipDeviceSynchronize() == 0
        out = np.zeros(SHAPE, order="F")
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def write(self, ptr, values):
        v = np.asfortranarray(values)
        assert self.hip.hipMemcpy(ptr, v.ctypes.data, v.nbytes, 1) == 0  # host to device
        assert self.hip.hipDeviceSynchronize() == 0


def boxes_of(m):
    """A list of disjoint boxes covering mask m (one per run of cells along the first axis)."""
    out = []
    for j in range(SHAPE[1]):
        i = 0
        while i < SHAPE[0]:
            if m[i, j]:
                k = i
                while k < SHAPE[0] and m[k, j]:
                    k += 1
                out.append(((i, j), (k, j + 1)))
                i = k
            else:
                i += 1
    return out


def run_sequence(scope, lib, seed, steps=250, memory=HostMemory):
    rng = np.random.default_rng(seed)
    initialized = bool(rng.integers(0, 2))
    a = np.asfortranarray(rng.integers(0, 100, SHAPE).astype(np.float64))
    model = Model(a, initialized)
    statuses = {"ARGUMENT": scope.ARGUMENT, "UNINITIALIZED": scope.UNINITIALIZED}
    counter = [1000.0]

    def fresh(cells):
        n = int(cells.sum())
        v = counter[0] + np.arange(n)
        counter[0] += n
        return v

    with scope.Domain(lib) as dom:
        h = dom.bind(a, initialized=initialized)
        for step in range(steps):
            op = rng.choice(["device", "device", "host", "host", "original", "observe", "forget", "discard",
                             "admit"], p=[0.22, 0.08, 0.2, 0.07, 0.13, 0.12, 0.08, 0.03, 0.07])
            what = f"seed {seed} step {step} {op}"
            if op in ("device", "host", "original"):
                r = mask(random_box(rng)) if rng.random() < 0.7 else np.zeros(SHAPE, bool)
                wbox = random_box(rng)
                w = mask(wbox)
                choice = rng.random()
                if choice < 0.35:
                    o = w.copy()
                    obox = wbox
                elif choice < 0.7:
                    obox = random_box(rng)
                    # a sub-box of the write box (or a box beyond it, now and then: an argument error)
                    if rng.random() < 0.9:
                        obox = (tuple(max(x, y) for x, y in zip(obox[0], wbox[0])),
                                tuple(max(min(x, y), max(p, q)) for x, y, p, q in
                                      zip(obox[1], wbox[1], obox[0], wbox[0])))
                    o = mask(obox)
                else:
                    o = np.zeros(SHAPE, bool)
                    obox = None
                rb = boxes_of(r)
                ob = [obox] if obox is not None else []
                access, keep = scope.make_access(0, rb, [wbox], ob)
                device = op == "device"
                expect = model.plan(r, w, o, device, strict=op != "original")
                if op == "original":
                    binding = (scope.PlanBinding * 1)(scope.PlanBinding(h, access))
                    st = lib.fort_scope_original_host_begin_v1(dom.ctx, binding, 1)
                elif device:
                    ptr = c_void_p()
                    st = lib.fort_scope_device_begin(dom.ctx, h, byref(access), byref(ptr))
                else:
                    st = lib.fort_scope_host_begin(dom.ctx, h, byref(access))
                if isinstance(expect, str):
                    assert st == statuses[expect], what
                    continue
                assert st == 0, (what, st)
                model.copy(expect, device)
                if device:
                    view = memory.read(ptr.value).copy()
                    seen = r & model.dc
                    assert np.array_equal(view[seen], model.dev[seen]), what
                    values = fresh(w)
                    view[w] = values
                    memory.write(ptr.value, view)
                    model.dev[w] = values
                    assert lib.fort_scope_device_end(dom.ctx, h) == 0, what
                    model.commit(w, o, True)
                else:
                    assert np.array_equal(a, model.host), what
                    target = o if op == "original" else w
                    values = fresh(target)
                    a[target] = values
                    model.host[target] = values
                    assert lib.fort_scope_host_end(dom.ctx, h) == 0, what
                    model.commit(w, o, False, original=op == "original")
            elif op == "observe":
                dom.observe(h)
                model.copy(model.physical() & ~model.hc, False)
            elif op == "forget":
                assert lib.fort_scope_forget_definition(dom.ctx, h) == 0, what
                model.init[:] = False
            elif op == "discard":
                assert lib.fort_scope_discard_v1(dom.ctx, h) == 0, what
                model.init[:] = model.hc[:] = model.dc[:] = False
            else:
                box = random_box(rng)
                m = mask(box)
                secs = scope.Sections([box])
                assert lib.fort_scope_admit_source_reads_v1(dom.ctx, h, secs.array, secs.count) == 0, what
                missing = m & ~model.init
                model.init |= missing
                model.hc |= missing & ~model.dc
            s = dom.stats()
            assert (s.upload_bytes, s.download_bytes) == (model.up, model.down), what
            assert np.array_equal(a, model.host), what
        final = model.physical() & ~model.hc
    # Closing publishes every physically held value the host lacks.
    model.host[final] = model.dev[final]
    assert np.array_equal(a, model.host), f"seed {seed} close"


@pytest.mark.parametrize("seed", range(12))
def test_random_access_sequences_follow_the_coherence_model(scope, cpu_runtime, seed):
    run_sequence(scope, cpu_runtime, seed)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_random_access_sequences_follow_the_coherence_model_on_the_device(scope, hip, seed):
    # The same sequences on libftc_hip.so: staged pitched copies, dense-span downloads and linear runs
    # must move exactly the model's cells.
    run_sequence(scope, hip, 100 + seed, steps=150, memory=DeviceMemory())


@pytest.mark.gpu
def test_fragmentation_reconciles_on_the_device(scope, hip):
    # compiler/tests/test_scoped_runtime.py:293-315 on libftc_hip.so: the reconciliation's copies go through
    # the device's copy paths (staged and linear), and closing publishes every value.
    size = 4 * 128
    memory = DeviceMemory()
    a = np.arange(float(size))
    expected = a.copy()
    with scope.Domain(hip) as dom:
        h = dom.bind(a)
        for i in range(2 * 128 + 16):
            box = ((i,), (i + 1,))
            access, _keep = scope.make_access(0, (), [box], [box])
            if i % 2:
                assert hip.fort_scope_host_begin(dom.ctx, h, byref(access)) == 0
                a[i] = -100.0 - i
                assert hip.fort_scope_host_end(dom.ctx, h) == 0
            else:
                ptr = c_void_p()
                assert hip.fort_scope_device_begin(dom.ctx, h, byref(access), byref(ptr)) == 0
                one = np.array([1000.0 + i])
                assert memory.hip.hipMemcpy(ptr.value + 8 * i, one.ctypes.data, 8, 1) == 0
                assert memory.hip.hipDeviceSynchronize() == 0
                assert hip.fort_scope_device_end(dom.ctx, h) == 0
            expected[i] = a[i] if i % 2 else 1000.0 + i
        assert dom.stats().reconciliations > 0
    assert np.array_equal(a, expected)
