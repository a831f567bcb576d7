"""The nests around the pressure solve of one domain (include/ftc_poisson.h: cpp_fort_shared_pressure_divergence,
cpp_fort_shared_pressure_projection and cpp_fort_shared_pressure_fix_reference, ELMM's prepoisson,
postpoisson and fixpressuretoreference) against the oracle (oracle/step.c orc_divergence,
orc_projection; oracle/reduce.c for the reference-plane sum), bit for bit, halos included; and a whole resident step
in ELMM's order (LES step, divergence, pressure solve, projection, all on the device in one epoch)
against the oracle's projected step (oracle/binding.py projected_step, whose solve is the FP64 direct
one of oracle/poisson.py): the FP64 provider within 1e-11 and the mixed one within 1e-10 of max|field|.
"""
import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu

SIZES = [(40, 24, 12), (33, 17, 9), (64, 64, 32)]


def with_ghosts(inner, ghosts):
    out = np.zeros(tuple(n + 2 * ghosts for n in inner.shape), order="F")
    out[ghosts:-ghosts, ghosts:-ghosts, ghosts:-ghosts] = inner
    return out


@pytest.mark.parametrize("n", SIZES)
def test_divergence_matches_the_oracle_bit_for_bit(ftc, scope, hip, oracle, fields, n):
    nx, ny, nz = n
    dx, dt = 10.0, 0.2
    U, V, W, _ = fields.les_fields(nx, ny, nz)
    rhs = np.full((nx + 2, ny + 2, nz + 2), 7.0, order="F")  # the ghosts are not written
    with scope.Domain(hip) as dom:
        hr = dom.bind(rhs, initialized=False)
        hu, hv, hw = (dom.bind(x.copy(order="F")) for x in (U, V, W))
        scope.check(hip, hip.cpp_fort_shared_pressure_divergence(
            dom.ctx, 1, hr, hu, hv, hw, ftc.dp(dx), ftc.dp(dx), ftc.dp(dx), ftc.dp(dt),
            ftc.ip(nx), ftc.ip(ny), ftc.ip(nz)))
        dom.flush()
        dom.observe(hr)
    want = oracle.divergence(U, V, W, dx, dx, dx, dt)
    assert bits_equal(rhs[1:-1, 1:-1, 1:-1], want)
    ghosts = np.ones(rhs.shape, bool)
    ghosts[1:-1, 1:-1, 1:-1] = False
    assert (rhs[ghosts] == 7.0).all()


@pytest.mark.parametrize("n", SIZES)
def test_projection_matches_the_oracle_bit_for_bit_including_halos(ftc, scope, hip, oracle, fields, n):
    nx, ny, nz = n
    dx, dt = 10.0, 0.2
    U, V, W, _ = fields.les_fields(nx, ny, nz, seed=5)
    phi = np.asfortranarray(np.random.default_rng(sum(n)).uniform(-1, 1, n))
    got = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F")]
    with scope.Domain(hip) as dom:
        hs = [dom.bind(x) for x in got]
        hp = dom.bind(with_ghosts(phi, 2))
        scope.check(hip, hip.cpp_fort_shared_pressure_projection(
            dom.ctx, 1, *hs, hp, ftc.dp(dx), ftc.dp(dx), ftc.dp(dx), ftc.dp(dt), ftc.ip(nx), ftc.ip(ny), ftc.ip(nz)))
        dom.flush()
        dom.observe(*hs)
    want = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F")]
    oracle.projection(*want, phi, dx, dx, dx, dt)
    for name, g, w in zip("UVW", got, want):
        assert bits_equal(g, w), name


def test_pressure_step_entries_refuse_wrong_layouts(ftc, scope, hip):
    a = np.zeros((6, 5, 4), order="F")
    b = np.zeros((7, 5, 4), order="F")
    with scope.Domain(hip) as dom:
        h = [dom.bind(x.copy(order="F")) for x in (a, a, a, a)]
        hb = dom.bind(b)
        args = (ftc.dp(1.0), ftc.dp(1.0), ftc.dp(1.0), ftc.dp(0.1), ftc.ip(4), ftc.ip(3), ftc.ip(2))
        assert hip.cpp_fort_shared_pressure_divergence(dom.ctx, 1, hb, h[1], h[2], h[3], *args) == scope.BOUNDARY
        assert hip.cpp_fort_shared_pressure_projection(dom.ctx, 1, h[0], h[1], h[2], h[3], *args) == scope.BOUNDARY
        assert hip.cpp_fort_shared_pressure_divergence(dom.ctx, 1, h[0], h[0], h[2], h[3], *args) == scope.ALIAS
        assert hip.cpp_fort_shared_pressure_divergence(dom.ctx, 0, *h, *args) == scope.DECLINED


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("n", [(48, 40, 24), (37, 29, 17)])
def test_a_resident_projected_step_matches_the_oracle(ftc, scope, hip, oracle, fields, mixed, n):
    from __graft_entry__ import _module

    pressure = _module("pressure")
    nx, ny, nz = n
    dx, dt, cs = 10.0, 0.2, 0.17
    U, V, W, S = fields.les_fields(nx, ny, nz, seed=9)
    src = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F"), S.copy(order="F")]
    dst = [np.zeros_like(U, order="F") for _ in range(4)]
    rhs = np.zeros_like(U, order="F")
    phi = np.zeros((nx + 4, ny + 4, nz + 4), order="F")
    sc = (ftc.dp(dx), ftc.dp(dx), ftc.dp(dx), ftc.dp(dt))
    with scope.Domain(hip) as dom:
        hs = [dom.bind(x) for x in src]
        hd = [dom.bind(x, initialized=True) for x in dst]
        hr = dom.bind(rhs, lower=(0, 0, 0), initialized=False)
        hp = dom.bind(phi, lower=(-1, -1, -1))
        solver = pressure.PressureSolver(hip, dom, n, (nx * dx, ny * dx, nz * dx), mixed)
        scope.check(hip, hip.cpp_fort_shared_les_step(dom.ctx, 1, *hd, *hs, *sc, ftc.dp(cs), ftc.ip(nx), ftc.ip(ny),
                                                       ftc.ip(nz), ftc.ip(1)))
        scope.check(hip, hip.cpp_fort_shared_pressure_divergence(dom.ctx, 1, hr, *hd[:3], *sc, ftc.ip(nx), ftc.ip(ny),
                                                                 ftc.ip(nz)))
        ext = solver.solve(hp, hr)
        assert ext.started == 1 and ext.committed == 1
        scope.check(hip, hip.cpp_fort_shared_pressure_projection(dom.ctx, 1, *hd[:3], hp, *sc, ftc.ip(nx),
                                                                 ftc.ip(ny), ftc.ip(nz)))
        dom.flush()
        dom.observe(*hd)
        stats = dom.stats()
        solver.destroy()
    want = oracle.projected_step(U, V, W, S, dx, dx, dx, dt, cs)
    tol = 1e-10 if mixed else 1e-11
    for name, g, w in zip("UVW", dst[:3], want[:3]):
        err = np.abs(g - w).max() / np.abs(w).max()
        assert err <= tol, (name, err)
    assert bits_equal(dst[3], want[3])  # the scalar is not projected
    # Everything stayed on the device between the entries: only the inputs went up.
    assert stats.upload_bytes <= 8 * (4 * U.size + dst[0].size * 4 + phi.size)


@pytest.mark.parametrize("n", [(40, 24, 12), (33, 17, 9)])
def test_fix_to_the_reference_plane_mean(ftc, scope, hip, oracle, n):
    # ELMM's fixpressuretoreference (elmm.toml:274-293): the mean of Pr(1:nx, 1:ny, k_ref) as a fixed-order
    # sum on the device, then phi -= that mean on the interior cells (the ghost layers keep their values),
    # bit for bit against the oracle's sum.
    import ctypes

    nx, ny, nz = n
    phi = np.asfortranarray(np.random.default_rng(nx).uniform(-1, 1, (nx + 4, ny + 4, nz + 4)))
    k_ref = nz // 2 + 1  # 1-based interior plane; the array index is k_ref + 1 with two ghosts below
    lo, hi = (2, 2, k_ref + 1), (nx + 2, ny + 2, k_ref + 2)
    got = phi.copy(order="F")
    total = ctypes.c_double(0.0)
    with scope.Domain(hip) as dom:
        hp = dom.bind(got, lower=(-1, -1, -1))
        scope.check(hip, hip.cpp_fort_shared_fixed_order_reduction(
            dom.ctx, 1, hp, ftc.ip(0), (ctypes.c_longlong * 3)(*lo), (ctypes.c_longlong * 3)(*hi), ctypes.byref(total)))
        ref = total.value / (nx * ny)
        scope.check(hip, hip.cpp_fort_shared_pressure_fix_reference(dom.ctx, 1, hp, ftc.dp(ref), ftc.ip(nx),
                                                                    ftc.ip(ny), ftc.ip(nz)))
        dom.flush()
        dom.observe(hp)
    want_total = oracle.fixed_order_reduce(phi, lo, hi, 0, 0.0)
    assert bits_equal(np.array([total.value]), np.array([want_total]))
    want = phi.copy(order="F")
    want[2:-2, 2:-2, 2:-2] -= want_total / (nx * ny)
    assert bits_equal(got, want)


K = lambda n: n[2] // 2 + 1  # the reference plane, 1-based interior index


def plane_box(n):
    nx, ny, nz = n
    return (2, 2, K(n) + 1), (nx + 2, ny + 2, K(n) + 2)


def test_fix_accepts_a_phi_the_solve_alone_defines(ftc, scope, hip, oracle):
    # phi bound as undefined: the solve defines its interior, and the plane sum and the shift touch only
    # that interior, so neither is refused as a read of undefined values.
    import ctypes
    from __graft_entry__ import _module
    from oracle import poisson as poisson_oracle

    pressure = _module("pressure")
    n = (32, 16, 8)
    nx, ny, nz = n
    h = (10.0, 10.0, 10.0)
    b = poisson_oracle.divergence_rhs(*n, h, seed=3)
    rhs = with_ghosts(b, 1)
    phi = np.zeros((nx + 4, ny + 4, nz + 4), order="F")
    lo, hi = plane_box(n)
    total = ctypes.c_double(0.0)
    with scope.Domain(hip) as dom:
        hr = dom.bind(rhs, lower=(0, 0, 0))
        hp = dom.bind(phi, lower=(-1, -1, -1), initialized=False)
        solver = pressure.PressureSolver(hip, dom, n, (nx * h[0], ny * h[1], nz * h[2]))
        solver.refinement(1e-12, 1, False)
        solver.solve(hp, hr)
        scope.check(hip, hip.cpp_fort_shared_fixed_order_reduction(
            dom.ctx, 1, hp, ftc.ip(0), (ctypes.c_longlong * 3)(*lo), (ctypes.c_longlong * 3)(*hi), ctypes.byref(total)))
        scope.check(hip, hip.cpp_fort_shared_pressure_fix_reference(dom.ctx, 1, hp, ftc.dp(total.value / (nx * ny)),
                                                                    ftc.ip(nx), ftc.ip(ny), ftc.ip(nz)))
        dom.flush()
        dom.observe(hp)
        solver.destroy()
    want = poisson_oracle.solve(b, h)
    want = want - want[:, :, K(n) - 1].mean()
    got = phi[2:-2, 2:-2, 2:-2]
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


def resident_steps(ftc, scope, hip, fields, n, steps, device_sum):
    """`steps` resident steps in ELMM's order: LES step, divergence, FP64 solve (one direct solve, no
    norms), projection, and fixpressuretoreference: the plane sum either back on the host (the plain
    reduction entry) or left on the device for the queued shift (device_sum). Returns the final U, V, W,
    phi's interior, and the domain's graph and queue statistics."""
    import ctypes
    from __graft_entry__ import _module

    pressure = _module("pressure")
    nx, ny, nz = n
    dx, dt, cs = 10.0, 0.2, 0.17
    start = fields.les_fields(nx, ny, nz, seed=21)
    a = [x.copy(order="F") for x in start]
    b = [np.zeros_like(start[0], order="F") for _ in range(4)]
    rhs = np.zeros_like(start[0], order="F")
    phi = np.zeros((nx + 4, ny + 4, nz + 4), order="F")
    partial = np.zeros(65536 + 256)
    ref = np.zeros(1)
    sc = (ftc.dp(dx), ftc.dp(dx), ftc.dp(dx), ftc.dp(dt))
    ext = (ftc.ip(nx), ftc.ip(ny), ftc.ip(nz))
    lo, hi = ((ctypes.c_longlong * 3)(*x) for x in plane_box(n))
    with scope.Domain(hip) as dom:
        ha, hb = [dom.bind(x) for x in a], [dom.bind(x, initialized=True) for x in b]
        hr, hp = dom.bind(rhs, lower=(0, 0, 0), initialized=False), dom.bind(phi, lower=(-1, -1, -1), initialized=False)
        hpart, href = dom.bind(partial, initialized=False), dom.bind(ref, initialized=False)
        solver = pressure.PressureSolver(hip, dom, n, (nx * dx, ny * dx, nz * dx))
        solver.refinement(1e-12, 1, False)
        for k in range(steps):
            src, dst = (ha, hb) if k % 2 == 0 else (hb, ha)
            scope.check(hip, hip.cpp_fort_shared_les_step(dom.ctx, 1, *dst, *src, *sc, ftc.dp(cs), *ext, ftc.ip(1)))
            scope.check(hip, hip.cpp_fort_shared_pressure_divergence(dom.ctx, 1, hr, *dst[:3], *sc, *ext))
            report = solver.solve(hp, hr)
            assert report.started == 1 and report.committed == 1
            scope.check(hip, hip.cpp_fort_shared_pressure_projection_bc(dom.ctx, 1, *dst[:3], hp, *sc, *ext,
                                                                        ftc.ip(0), ftc.ip(0)))
            if device_sum:
                scope.check(hip, hip.cpp_fort_shared_fixed_order_reduction_device(dom.ctx, 1, hp, hpart, href, ftc.ip(0),
                                                                                  lo, hi, ftc.dp(0.0)))
                scope.check(hip, hip.cpp_fort_shared_pressure_fix_reference_sum(dom.ctx, 1, hp, href, *ext))
            else:
                total = ctypes.c_double(0.0)
                scope.check(hip, hip.cpp_fort_shared_fixed_order_reduction(dom.ctx, 1, hp, ftc.ip(0), lo, hi,
                                                                           ctypes.byref(total)))
                scope.check(hip, hip.cpp_fort_shared_pressure_fix_reference(dom.ctx, 1, hp, ftc.dp(total.value / (nx * ny)),
                                                                            *ext))
            dom.flush()
        final = dst
        dom.observe(*final[:3], hp)
        gstats, qstats = dom.graph_stats(), dom.queue_stats()
        solver.destroy()
    out = (b if steps % 2 == 1 else a)[:3]
    return [x.copy() for x in out], phi[2:-2, 2:-2, 2:-2].copy(), gstats, qstats


@pytest.mark.parametrize("n", [(48, 40, 24), (288, 256, 8)])
def test_resident_steps_with_the_reference_fix_match_the_oracle(ftc, scope, hip, oracle, fields, n):
    # Six steps, with a host round trip of the plane sum and a shift argument that changes every step, so
    # the shift's graph node takes parameter updates; (288, 256, 8) has nx*ny > 65536 cells in the plane,
    # so its sum takes the two-level (slots, then folds) path.
    print(f"resident steps with the reference fix {n}", flush=True)
    nx, ny, nz = n
    steps = 6
    got, phi, _, _ = resident_steps(ftc, scope, hip, fields, n, steps, device_sum=False)
    state = list(fields.les_fields(nx, ny, nz, seed=21))
    for _ in range(steps):
        *state, want_phi = oracle.projected_step(*state, 10.0, 10.0, 10.0, 0.2, 0.17)
    lo, hi = plane_box(n)
    want_phi = want_phi - oracle.fixed_order_reduce(with_ghosts(want_phi, 2), lo, hi, 0, 0.0) / (nx * ny)
    for name, g, w in zip("UVW", got, state[:3]):
        assert np.abs(g - w).max() <= 1e-11 * np.abs(w).max(), name
    assert np.abs(phi - want_phi).max() <= 1e-11 * np.abs(want_phi).max()


@pytest.mark.parametrize("n", [(64, 32, 16), (288, 256, 8)])
def test_a_step_held_for_one_graph_replays_bit_identical_to_the_host_path(ftc, scope, hip, fields, monkeypatch, n):
    # With FORT_SCOPE_GRAPH_QUEUE_WORK=0 the queue holds a whole step: the LES step, the divergence, the
    # provider's solve (captured from its stream into one node), the projection, the queued plane sum and
    # the shift that reads it on the device. Each flush submits one chain graph: the first sighting
    # launches directly, the two ping-pong variants are instantiated once, and every later step replays
    # one of them, refreshing only the captured solve. The fields equal the host-sum path bit for bit.
    print(f"whole-step graph {n}", flush=True)
    steps = 7
    host = resident_steps(ftc, scope, hip, fields, n, steps, device_sum=False)
    monkeypatch.setenv("FORT_SCOPE_GRAPH_QUEUE_WORK", "0")
    dev = resident_steps(ftc, scope, hip, fields, n, steps, device_sum=True)
    for g, w in zip(dev[0] + [dev[1]], host[0] + [host[1]]):
        assert bits_equal(g, w)
    gstats, qstats = dev[2], dev[3]
    assert gstats.submissions == steps
    assert gstats.instantiations == 2
    assert gstats.replays == steps - 3
    assert gstats.parameter_updates == steps - 3  # the captured solve, once per replay


def test_the_projection_refuses_walls_in_x_or_y(ftc, scope, hip):
    a = np.zeros((6, 5, 4), order="F")
    p = np.zeros((8, 7, 6), order="F")
    with scope.Domain(hip) as dom:
        h = [dom.bind(a.copy(order="F")) for _ in range(3)]
        hp = dom.bind(p)
        args = (ftc.dp(1.0), ftc.dp(1.0), ftc.dp(1.0), ftc.dp(0.1), ftc.ip(4), ftc.ip(3), ftc.ip(2))
        for bx, by in ((4, 0), (0, 4), (4, 4)):
            assert hip.cpp_fort_shared_pressure_projection_bc(dom.ctx, 1, *h, hp, *args, ftc.ip(bx),
                                                              ftc.ip(by)) == scope.BOUNDARY
