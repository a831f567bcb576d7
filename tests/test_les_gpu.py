"""Parity of the HIP LES kernels with the oracle, through the C-ABI entries and the fort_scope runtime.

Every FP64 output is compared bit for bit: the kernels are compiled without contraction and keep the
operand order of the nests, so the device reproduces the native result exactly (halo and index handling
included: ghost cells are compared too).
"""
import numpy as np
import pytest

from conftest import bits_equal, describe_mismatch

pytestmark = pytest.mark.gpu

ENTRY = {"cdu": "cpp_fort_shared_momentumadvection_cdu", "cdv": "cpp_fort_shared_momentumadvection_cdv",
         "cdw": "cpp_fort_shared_momentumadvection_cdw"}


def run_component(ftc, scope, hip, name, U, V, W, d, loops):
    out = np.zeros_like(U, order="F")
    with scope.Domain(hip) as dom:
        h = [dom.bind(a) for a in (out, U, V, W)]
        status = getattr(hip, ENTRY[name])(dom.ctx, 1, *h, ftc.dp(d[0]), ftc.dp(d[1]), ftc.dp(d[2]),
                                           ftc.ip(loops[0]), ftc.ip(loops[1]), ftc.ip(loops[2]))
        scope.check(hip, status)
        dom.flush()
        dom.observe(h[0])
    return out


@pytest.mark.parametrize("name", ["cdu", "cdv", "cdw"])
@pytest.mark.parametrize("shape,loops", [((18, 18, 18), (16, 16, 16)), ((66, 66, 66), (64, 64, 64)),
                                         ((71, 9, 14), (69, 7, 12)), ((130, 20, 10), (120, 13, 8))])
def test_reference_advection_entries_are_bit_exact(ftc, scope, hip, oracle, name, shape, loops):
    rng = np.random.default_rng(sum(shape))
    U, V, W = (np.asfortranarray(rng.uniform(-1, 1, shape)) for _ in range(3))
    d = (0.37, 1.9, 0.61)
    got = run_component(ftc, scope, hip, name, U, V, W, d, loops)
    want = oracle.component(name, U, V, W, *d, *loops)
    n1, n2, n3 = loops
    interior = (slice(1, n1 + 1), slice(1, n2 + 1), slice(1, n3 + 1))
    assert bits_equal(got[interior], want[interior]), describe_mismatch(got[interior], want[interior])


def test_golden_stencil_dump_on_the_device(ftc, scope, hip):
    from pathlib import Path

    g = np.load(Path(__file__).resolve().parent / "golden" / "stencil_16.npz")
    U, V, W = (np.asfortranarray(g[k]) for k in "UVW")
    for name, key in (("cdu", "U2"), ("cdv", "V2"), ("cdw", "W2")):
        got = run_component(ftc, scope, hip, name, U, V, W, (1 / 16, 1 / 16, 1 / 16), (16, 16, 16))
        inner = np.ascontiguousarray(got[1:-1, 1:-1, 1:-1])
        assert bits_equal(inner, np.ascontiguousarray(g[key])), describe_mismatch(inner, g[key])


def tendency(ftc, scope, hip, U, V, W, S, d, cs):
    # d: one spacing for all three directions, or (dx, dy, dz)
    dx, dy, dz = d if isinstance(d, tuple) else (d, d, d)
    nx, ny, nz = (n - 2 for n in U.shape)
    outs = [np.zeros_like(U, order="F") for _ in range(4)]
    with scope.Domain(hip) as dom:
        hin = [dom.bind(a) for a in (U, V, W, S)]
        hout = [dom.bind(a) for a in outs]
        st = hip.cpp_fort_shared_les_tendency(dom.ctx, 1, *hout, *hin, ftc.dp(dx), ftc.dp(dy), ftc.dp(dz),
                                              ftc.dp(cs or 0.0), ftc.ip(nx), ftc.ip(ny), ftc.ip(nz),
                                              ftc.ip(0 if cs is None else 1))
        scope.check(hip, st)
        dom.flush()
        dom.observe(*hout)
    return outs


@pytest.mark.parametrize("cs", [None, 0.17])
@pytest.mark.parametrize("n", [(64, 64, 64), (70, 37, 19), (5, 3, 2), (130, 9, 33)])
def test_les_tendency_is_bit_exact(ftc, scope, hip, oracle, fields, cs, n):
    U, V, W, S = fields.les_fields(*n)
    got = tendency(ftc, scope, hip, U, V, W, S, 10.0, cs)
    want = oracle.les_tendency(U, V, W, S, 10.0, 10.0, 10.0, cs)
    nx, ny, nz = n
    for name, g, w, kz in zip("UVWS", got, want, (nz, nz, nz - 1, nz)):
        sl = (slice(1, nx + 1), slice(1, ny + 1), slice(1, kz + 1))
        assert bits_equal(g[sl], w[sl]), f"{name}: " + describe_mismatch(g[sl], w[sl])


@pytest.mark.parametrize("spacing", [(10.0, 8.0, 12.5), (4.0, 4.0, 2.5)])
@pytest.mark.parametrize("n", [(64, 64, 64), (70, 37, 19)])
def test_les_tendency_with_unequal_spacings_is_bit_exact(ftc, scope, hip, oracle, fields, spacing, n):
    # Unequal spacings take the kernel that holds every per-direction coefficient; equal ones (the tests
    # above) take the one that holds each once.
    U, V, W, S = fields.les_fields(*n)
    got = tendency(ftc, scope, hip, U, V, W, S, spacing, 0.17)
    want = oracle.les_tendency(U, V, W, S, *spacing, 0.17)
    nx, ny, nz = n
    for name, g, w, kz in zip("UVWS", got, want, (nz, nz, nz - 1, nz)):
        sl = (slice(1, nx + 1), slice(1, ny + 1), slice(1, kz + 1))
        assert bits_equal(g[sl], w[sl]), f"{name} {spacing}: " + describe_mismatch(g[sl], w[sl])


@pytest.mark.parametrize("n,steps", [((48, 40, 24), 6), ((33, 17, 9), 3), ((1, 2, 1), 2), ((512, 512, 128), 4)])
def test_les_steps_with_graph_replay_are_bit_exact_including_halos(ftc, scope, hip, oracle, fields, n, steps):
    nx, ny, nz = n
    print(f"LES steps {n} x {steps}", flush=True)
    dx, dt, cs = 10.0, 0.2, 0.17
    U, V, W, S = fields.les_fields(nx, ny, nz)
    a = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F"), S.copy(order="F")]
    b = [np.zeros_like(U, order="F") for _ in range(4)]
    with scope.Domain(hip) as dom:
        ha = [dom.bind(x) for x in a]
        hb = [dom.bind(x, initialized=True) for x in b]
        for step in range(steps):
            src, dst = (ha, hb) if step % 2 == 0 else (hb, ha)
            st = hip.cpp_fort_shared_les_step(dom.ctx, 1, *dst, *src, ftc.dp(dx), ftc.dp(dx), ftc.dp(dx),
                                              ftc.dp(dt), ftc.dp(cs), ftc.ip(nx), ftc.ip(ny), ftc.ip(nz), ftc.ip(1))
            scope.check(hip, st)
            dom.flush()
        final = hb if steps % 2 == 1 else ha
        dom.observe(*final)
        gstats = dom.graph_stats()
    got = b if steps % 2 == 1 else a
    want = [U, V, W, S]
    for _ in range(steps):
        want = oracle.les_step(*want, dx, dx, dx, dt, cs)
    for name, g, w in zip("UVWS", got, want):
        assert bits_equal(g, w), f"{name}: " + describe_mismatch(g, w)
    # One topology with two alternating argument sets: the first step launches directly, the next two
    # instantiate one graph each, and every later step replays one of them without a parameter update.
    if steps >= 3:
        assert gstats.instantiations == 2
        assert gstats.replays == steps - 3
        assert gstats.parameter_updates == 0


@pytest.mark.parametrize("variant", [{"FTC_LES_ROWS": "2"}, {"FTC_LES_XCD_ORDER": "1"}, {"FTC_LES_UNSTAGED": "1"},
                                     {"FTC_LES_KCHUNK": "5"}, {"FTC_LES_PERSISTENT": "1"},
                                     {"FTC_LES_PERSISTENT": "1", "FTC_LES_ROWS": "2"}, {"FTC_LES_WIDTH": "32"},
                                     {"FTC_LES_WIDTH": "128"}, {"FTC_LES_WIDTH": "32", "FTC_LES_PERSISTENT": "1"},
                                     {"FTC_LES_ANISO": "1"}, {"FTC_LES_ANISO": "1", "FTC_LES_WIDTH": "32"},
                                     {"FTC_LES_WIDTH": "32", "FTC_LES_TWO_BARRIERS": "1"},
                                     {"FTC_LES_WIDTH": "32", "FTC_LES_TWO_BARRIERS": "1", "FTC_LES_PERSISTENT": "1"},
                                     {"FTC_LES_WIDTH": "32", "FTC_LES_TWO_BARRIERS": "1", "FTC_LES_ANISO": "1"}])
@pytest.mark.parametrize("n", [(70, 37, 19), (130, 9, 33)])
def test_selectable_step_variants_are_bit_exact(ftc, scope, hip, oracle, fields, monkeypatch, variant, n):
    # The entry reads these switches at every call: two rows per thread, the XCD-contiguous tile order,
    # the unstaged kernel, a forced k chunk, the persistent banded walk (one resident round of
    # workgroups), the 32 x 8 and 128 x 2 tiles, the 32 x 8 tile with two barriers per plane, and the
    # general kernel at equal spacings. Each must give the default kernel's (the oracle's) bits.
    for key, value in variant.items():
        monkeypatch.setenv(key, value)
    U, V, W, S = fields.les_fields(*n)
    got = tendency(ftc, scope, hip, U, V, W, S, 10.0, 0.17)
    want = oracle.les_tendency(U, V, W, S, 10.0, 10.0, 10.0, 0.17)
    nx, ny, nz = n
    for name, g, w, kz in zip("UVWS", got, want, (nz, nz, nz - 1, nz)):
        sl = (slice(1, nx + 1), slice(1, ny + 1), slice(1, kz + 1))
        assert bits_equal(g[sl], w[sl]), f"{name} {variant}: " + describe_mismatch(g[sl], w[sl])


def les_steps(ftc, scope, hip, fields, n, steps, spacing=(10.0, 10.0, 10.0), dt=0.2, cs=0.17):
    """`steps` LES steps through the step entry from the seeded fields; returns the device's final fields."""
    nx, ny, nz = n
    U, V, W, S = fields.les_fields(nx, ny, nz, *spacing)
    a = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F"), S.copy(order="F")]
    b = [np.zeros_like(U, order="F") for _ in range(4)]
    with scope.Domain(hip) as dom:
        ha = [dom.bind(x) for x in a]
        hb = [dom.bind(x, initialized=True) for x in b]
        for step in range(steps):
            src, dst = (ha, hb) if step % 2 == 0 else (hb, ha)
            st = hip.cpp_fort_shared_les_step(dom.ctx, 1, *dst, *src, *(ftc.dp(d) for d in spacing), ftc.dp(dt),
                                              ftc.dp(cs), ftc.ip(nx), ftc.ip(ny), ftc.ip(nz), ftc.ip(1))
            scope.check(hip, st)
            dom.flush()
        final = hb if steps % 2 == 1 else ha
        dom.observe(*final)
    return (b if steps % 2 == 1 else a), (U, V, W, S)


@pytest.mark.parametrize("variant", [{}, {"FTC_LES_SHAPE": "64x8"}, {"FTC_LES_SHAPE": "64x4"},
                                     {"FTC_LES_KCHUNK": "6"}, {"FTC_LES_SHAPE": "64x8", "FTC_LES_KCHUNK": "12"},
                                     {"FTC_LES_SHAPE": "64x4", "FTC_LES_KCHUNK": "18"}, {"FTC_LES_LEGACY": "1"},
                                     {"FTC_LES_ANISO": "1"}, {"FTC_LES_ANISO": "1", "FTC_LES_SHAPE": "64x8"}])
@pytest.mark.parametrize("n", [(70, 37, 19), (130, 9, 33), (5, 3, 2), (64, 16, 13)])
def test_step_kernel_shapes_and_chunks_are_bit_exact_including_halos(ftc, scope, hip, oracle, fields, monkeypatch,
                                                                     variant, n):
    # The step entry's buffer-addressed kernel in each tile shape (64 x 8, 32 x 16, 64 x 4 columns), with k
    # chunks forced short (the last chunk taking the rest), in its general-coefficient form, and the les_sgs3
    # kernel it replaced: every one must give the oracle's bits, ghost cells included, over three steps.
    for key, value in variant.items():
        monkeypatch.setenv(key, value)
    print(f"LES step kernel {variant} {n}", flush=True)
    got, start = les_steps(ftc, scope, hip, fields, n, 3)
    want = list(start)
    for _ in range(3):
        want = oracle.les_step(*want, 10.0, 10.0, 10.0, 0.2, 0.17)
    for name, g, w in zip("UVWS", got, want):
        assert bits_equal(g, w), f"{name} {variant}: " + describe_mismatch(g, w)


@pytest.mark.parametrize("spacing", [(10.0, 8.0, 12.5), (4.0, 4.0, 2.5)])
@pytest.mark.parametrize("n", [(70, 37, 19), (33, 17, 9)])
def test_steps_with_unequal_spacings_are_bit_exact_including_halos(ftc, scope, hip, oracle, fields, spacing, n):
    got, start = les_steps(ftc, scope, hip, fields, n, 2, spacing)
    want = list(start)
    for _ in range(2):
        want = oracle.les_step(*want, *spacing, 0.2, 0.17)
    for name, g, w in zip("UVWS", got, want):
        assert bits_equal(g, w), f"{name} {spacing}: " + describe_mismatch(g, w)


def test_hundred_resident_steps_at_full_size_equal_the_native_build(ftc, scope, hip, fields, tmp_path):
    # 512 x 512 x 128 fields resident for 100 steps, each a graph replay, against the native Fortran build of
    # the same step (oracle/_build/les_native, itself pinned to the C restatement) run for 100 steps from the
    # same fields: every value, ghost cells included, bit for bit.
    import os
    import subprocess
    from pathlib import Path

    native = Path(__file__).resolve().parents[1] / "oracle" / "_build" / "les_native"
    assert native.exists(), "oracle/_build/les_native is not built (make -C oracle)"
    nx, ny, nz, steps = 512, 512, 128, 100
    print(f"LES {steps} steps at {nx}x{ny}x{nz} against the native build", flush=True)
    got, start = les_steps(ftc, scope, hip, fields, (nx, ny, nz), steps)
    np.concatenate([a.ravel(order="F") for a in start]).tofile(tmp_path / "in.bin")
    threads = str(min(16, len(os.sched_getaffinity(0))))
    run = subprocess.run([str(native), str(nx), str(ny), str(nz), str(steps - 1), "1", str(tmp_path / "in.bin"),
                          str(tmp_path / "first.bin"), str(tmp_path / "final.bin")], capture_output=True, text=True,
                         timeout=240, env=dict(os.environ, OMP_NUM_THREADS=threads))
    assert run.returncode == 0, run.stderr
    raw = np.fromfile(tmp_path / "final.bin", dtype=np.float64)
    size = (nx + 2) * (ny + 2) * (nz + 2)
    for f, (name, g) in enumerate(zip("UVWS", got)):
        want = raw[f * size:(f + 1) * size].reshape((nx + 2, ny + 2, nz + 2), order="F")
        assert bits_equal(g, want), f"{name}: " + describe_mismatch(g, want)
