"""MFMA lane layouts, the contraction engine and the chain contraction D = A*B*C on the device.

The lane layouts the kernels assume are checked first with one-hot operands (the probe runs one MFMA,
in bounds by construction). Products are then compared with numpy's FP64 matmul: the FP64 MFMA path to
rounding (1e-13 normwise), the split-bf16 path to its stated error (1e-4 normwise).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def probe(hip, kind, a_lane, a_elem, b_lane, b_elem):
    out = (ctypes.c_double * 256)()
    assert hip.ftc_mfma_layout_probe(kind, a_lane, a_elem, b_lane, b_elem, out) == 0
    return np.array(out).reshape(64, 4)


def test_f64_mfma_lane_layout(hip):
    # The FP64 accumulator interleaves rows across its four registers: element r of lane l is row l/16 + 4r.
    lanes = np.arange(64)[:, None]
    regs = np.arange(4)[None, :]
    out_row = lanes // 16 + 4 * regs
    out_col = np.broadcast_to(lanes % 16, (64, 4))
    for a in range(64):
        got = probe(hip, 0, a, 0, -1, 0)
        np.testing.assert_array_equal(got, (out_row == a % 16).astype(float), err_msg=f"A lane {a}")
    for b in range(64):
        got = probe(hip, 0, -1, 0, b, 0)
        np.testing.assert_array_equal(got, (out_col == b % 16).astype(float), err_msg=f"B lane {b}")
    for a, b in ((0, 0), (5, 21), (17, 33), (40, 40), (63, 15), (50, 60)):
        assert probe(hip, 0, a, 0, b, 0).sum() == (1.0 if a // 16 == b // 16 else 0.0)


def test_bf16_mfma_lane_layout(hip):
    lanes = np.arange(64)[:, None]
    regs = np.arange(4)[None, :]
    out_row = 4 * (lanes // 16) + regs
    out_col = np.broadcast_to(lanes % 16, (64, 4))
    for a in range(0, 64, 3):
        for e in (0, 3, 7):
            got = probe(hip, 1, a, e, -1, 0)
            np.testing.assert_array_equal(got, (out_row == a % 16).astype(float), err_msg=f"A lane {a} elem {e}")
    for b in range(0, 64, 5):
        got = probe(hip, 1, -1, 0, b, 2)
        np.testing.assert_array_equal(got, (out_col == b % 16).astype(float), err_msg=f"B lane {b}")
    k = lambda lane, elem: 8 * (lane // 16) + elem
    for (a, ea), (b, eb) in (((0, 0), (0, 0)), ((3, 5), (20, 5)), ((17, 2), (49, 2)), ((33, 7), (34, 6)),
                             ((60, 1), (44, 1)), ((10, 4), (26, 4))):
        assert probe(hip, 1, a, ea, b, eb).sum() == (1.0 if k(a, ea) == k(b, eb) else 0.0)


def device_gemm(ftc, scope, hip, A, B, C, precision, batch, tile=-1, alpha=1.0, beta=0.0, m=None, n=None, k=None,
                flags=0):
    """Runs ftc_gemm_device on registered arrays (A: lda x K x batch, etc.) and returns C."""
    lda, ldb, ldc = A.shape[0], B.shape[0], C.shape[0]
    m = m or A.shape[0]
    k = k or A.shape[1]
    n = n or B.shape[1]
    types = {np.dtype(np.float64): 0, np.dtype(np.float32): 1}
    with scope.Domain(hip) as dom:
        hs = [dom.bind(X) for X in (A, B, C)]
        binds = []
        keep = []
        output = scope.READ_ALL | scope.WRITE_ALL | scope.OVERWRITE_ALL
        for h, access in zip(hs, (scope.READ_ALL, scope.READ_ALL, output)):
            acc, kp = scope.make_access(access)
            keep.append(kp)
            binds.append(scope.PlanBinding(h, acc))
        arr = (scope.PlanBinding * 3)(*binds)
        devs = (ctypes.c_void_p * 3)()
        scope.check(hip, hip.fort_scope_access_begin_v1(dom.ctx, arr, 3, 1, devs))
        desc = ftc.GemmDesc(devs[0], devs[1], devs[2], types[A.dtype], types[B.dtype], types[C.dtype], m, n, k,
                            batch, lda, ldb, ldc, A.shape[0] * A.shape[1] if A.ndim == 3 else 0,
                            B.shape[0] * B.shape[1] if B.ndim == 3 else 0, C.shape[0] * C.shape[1], alpha, beta,
                            flags)
        stream = ctypes.c_void_p(dom.stream())
        scope.check(hip, hip.ftc_gemm_device(ctypes.byref(desc), precision, tile, stream))
        for h in hs:
            scope.check(hip, hip.fort_scope_device_end(dom.ctx, h))
        dom.observe(hs[2])
    return C


@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 1e-4)])
@pytest.mark.parametrize("m,n,k,batch", [(128, 128, 128, 1), (100, 37, 53, 3), (513, 260, 96, 1), (32, 128, 33, 5)])
def test_contraction_matches_numpy(ftc, scope, hip, precision, tol, m, n, k, batch):
    rng = np.random.default_rng(m * n + k)
    A = np.asfortranarray(rng.uniform(-1, 1, (m + 3, k, batch)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n, batch)))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (m + 1, n, batch)))
    C = device_gemm(ftc, scope, hip, A, B, C0.copy(order="F"), precision, batch, alpha=0.75, beta=-0.5, m=m, n=n, k=k)
    for b in range(batch):
        want = 0.75 * (A[:m, :, b] @ B[:, :, b]) - 0.5 * C0[:m, :, b]
        err = np.abs(C[:m, :, b] - want).max() / np.abs(want).max()
        assert err <= tol, (b, err)
        assert np.array_equal(C[m:, :, b], C0[m:, :, b])  # rows past M untouched


@pytest.mark.parametrize("tile", [0, 1, 2])
def test_every_tile_shape_gives_the_same_fp64_product(ftc, scope, hip, tile):
    rng = np.random.default_rng(tile)
    A = np.asfortranarray(rng.uniform(-1, 1, (96, 80)))
    B = np.asfortranarray(rng.uniform(-1, 1, (80, 72)))
    C = device_gemm(ftc, scope, hip, A, B, np.zeros((96, 72), order="F"), 0, 1, tile=tile)
    assert np.abs(C - A @ B).max() <= 1e-13 * np.abs(A @ B).max()


def test_float_operands_and_outputs(ftc, scope, hip):
    rng = np.random.default_rng(9)
    A = np.asfortranarray(rng.uniform(-1, 1, (70, 50)).astype(np.float32))
    B = np.asfortranarray(rng.uniform(-1, 1, (50, 40)).astype(np.float32))
    C = device_gemm(ftc, scope, hip, A, B, np.zeros((70, 40), order="F"), 1, 1)
    want = A.astype(np.float64) @ B.astype(np.float64)
    assert np.abs(C - want).max() <= 1e-4 * np.abs(want).max()


F, D = np.float32, np.float64


@pytest.mark.parametrize("types,tout", [((F, D, F), True), ((F, F, F), True), ((F, F, F), False),
                                        ((F, F, D), False), ((D, D, D), False)])
@pytest.mark.parametrize("m,n,k,batch,pad", [(128, 128, 128, 1, 0), (100, 37, 53, 3, 1), (64, 256, 96, 2, 8),
                                             (33, 70, 200, 1, 3)])
@pytest.mark.parametrize("tile", [-1, 0, 1, 2])
def test_k_contiguous_split_bf16(ftc, scope, hip, types, tout, m, n, k, batch, pad, tile):
    # A stored K-contiguous (A(m, k) at m*lda + k, i.e. the Fortran array At(k, m)); C optionally written
    # transposed (Ct(n, m)). pad > 0 gives leading dimensions that break the 16-byte vector path.
    rng = np.random.default_rng(m + 7 * n + 13 * k + pad)
    At = np.asfortranarray(rng.uniform(-1, 1, (k + pad, m, batch)).astype(types[0]))
    B = np.asfortranarray(rng.uniform(-1, 1, (k + pad, n, batch)).astype(types[1]))
    shape = (n + pad, m, batch) if tout else (m + pad, n, batch)
    C0 = np.asfortranarray(rng.uniform(-1, 1, shape).astype(types[2]))
    flags = 1 | (2 if tout else 0)
    C = device_gemm(ftc, scope, hip, At, B, C0.copy(order="F"), 1, batch, tile=tile, alpha=1.25, beta=0.5, m=m, n=n,
                    k=k, flags=flags)
    for b in range(batch):
        want = 1.25 * (At[:k, :, b].astype(D).T @ B[:k, :, b].astype(D))
        got = C[:n, :, b].T if tout else C[:m, :, b]
        want = want + 0.5 * (C0[:n, :, b].T if tout else C0[:m, :, b]).astype(D)
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= 1e-4, (b, err)
        rest = (C[n:, :, b], C0[n:, :, b]) if tout else (C[m:, :, b], C0[m:, :, b])
        assert np.array_equal(*rest)  # padding past the logical extent untouched


def test_layout_flags_are_checked_before_any_launch(ftc, scope, hip):
    # Rejected on the host: layout flags with FP64, a transposed C without a K-contiguous A, unknown bits.
    A = np.zeros((8, 8), order="F")
    for precision, flags in ((0, 1), (1, 2), (1, 4), (1, 1 | 8)):
        with pytest.raises(RuntimeError):
            device_gemm(ftc, scope, hip, A, A.copy(order="F"), A.copy(order="F"), precision, 1, flags=flags)


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 2e-4)])
@pytest.mark.parametrize("n", [96, 301, 256])  # 256: whole 128 x 128 tiles (the pre-split K-contiguous path)
def test_chain_contraction_entry(ftc, scope, hip, precision, tol, n):
    rng = np.random.default_rng(n + precision)
    A, B, C = (np.asfortranarray(rng.uniform(-1, 1, (n, n))) for _ in range(3))
    D = np.zeros((n, n), order="F")
    with scope.Domain(hip) as dom:
        hd, ha, hb, hc = (dom.bind(X) for X in (D, A, B, C))
        scope.check(hip, hip.cpp_fort_shared_chain_contraction(dom.ctx, 1, hd, ha, hb, hc, ftc.ip(n),
                                                               ftc.ip(precision)))
        dom.flush()
        dom.observe(hd)
    want = (A @ B) @ C
    assert np.abs(D - want).max() / np.abs(want).max() <= tol


def test_small_integer_chain_is_exact_in_fp64(ftc, scope, hip):
    rng = np.random.default_rng(4)
    n = 64
    A, B, C = (np.asfortranarray(rng.integers(-3, 4, (n, n)).astype(np.float64)) for _ in range(3))
    D = np.zeros((n, n), order="F")
    with scope.Domain(hip) as dom:
        hd, ha, hb, hc = (dom.bind(X) for X in (D, A, B, C))
        scope.check(hip, hip.cpp_fort_shared_chain_contraction(dom.ctx, 1, hd, ha, hb, hc, ftc.ip(n), ftc.ip(0)))
        dom.flush()
        dom.observe(hd)
    assert np.array_equal(D, (A @ B) @ C)
