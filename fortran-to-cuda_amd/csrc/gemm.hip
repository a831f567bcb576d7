// MFMA contractions for gfx950 (see gemm.hpp for the operation and precisions), the dense chain
// contraction entry D = A*B*C, and an in-bounds probe of the MFMA lane layouts the kernels assume
// (the layouts are listed in gemm_kernels.hpp).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ftc_dense.h"
#include "entry.hpp"
#include "gemm.hpp"
#include "gemm_kernels.hpp"

namespace ftc {

// ------------------------------------------------------------------------------------------------ host
struct Tile {
    int bm, bn;
};
static const Tile kTiles[] = {{128, 128}, {64, 64}, {32, 32}};
constexpr int kTileCount = 3;

int gemm_tile_count(int) { return kTileCount; }

template <int BM, int BN>
static const void *f64_kernel()
{
    constexpr int WM = 2, WN = 2;
    // The 128 x 128 tile runs one workgroup per CU (368+ registers per lane), so its barriers are exposed:
    // two LDS stages of 32 elements of K (one barrier per stage) took the N = 4096 chain from 53 % to 56 %
    // of the FP64 MFMA peak and the Poisson transforms from 864 to 820 us.
    if constexpr (BM == 128 && BN == 128)
        return reinterpret_cast<const void *>(&gemm_f64_db_kernel<BM, BN, WM, WN, 32>);
    return reinterpret_cast<const void *>(&gemm_f64_kernel<BM, BN, WM, WN>);
}

template <class TA, class TB, class TO, int BM, int BN>
static const void *bf16_kernel()
{
    return reinterpret_cast<const void *>(&gemm_bf16x3_kernel<TA, TB, TO, BM, BN, 2, 2>);
}

template <class TA, class TB, class TO>
static const void *bf16_by_tile(int tile)
{
    switch (tile) {
    case 0: return bf16_kernel<TA, TB, TO, 128, 128>();
    case 1: return bf16_kernel<TA, TB, TO, 64, 64>();
    default: return bf16_kernel<TA, TB, TO, 32, 32>();
    }
}

template <class TA, class TB, class TO, bool TOUT>
static const void *tn_by_tile(int tile)
{
    // The 128 x 128 tile stages 64 elements of K per barrier pair (two MFMA K-slices): 5-10 % faster than 32
    // on the 512 x 512 x 128 Poisson transforms, at 72 KB of LDS (two workgroups per CU).
    switch (tile) {
    case 0: return reinterpret_cast<const void *>(&gemm_bf16x3_tn_kernel<TA, TB, TO, 128, 128, TOUT, 64>);
    case 1: return reinterpret_cast<const void *>(&gemm_bf16x3_tn_kernel<TA, TB, TO, 64, 64, TOUT>);
    default: return reinterpret_cast<const void *>(&gemm_bf16x3_tn_kernel<TA, TB, TO, 32, 32, TOUT>);
    }
}

int gemm_default_tile(const GemmCall &c, int precision)
{
    // The largest tile that still yields about two workgroups per CU of a 256-CU device; small
    // operands (the ensemble's subdomains) fall to the smaller tiles.
    (void)precision;
    for (int tile = 0; tile < kTileCount; ++tile) {
        const long long blocks = (long long)((c.M + kTiles[tile].bm - 1) / kTiles[tile].bm) *
                                 ((c.N + kTiles[tile].bn - 1) / kTiles[tile].bn) * c.batch;
        if (blocks >= 512 || tile == kTileCount - 1) return tile;
    }
    return kTileCount - 1;
}

static int select_kernel(const GemmCall &c, int precision, int tile, const void **fn, dim3 *grid)
{
    if (tile < 0 || tile >= kTileCount) tile = gemm_default_tile(c, precision);
    if (c.shift && !(precision == kBf16x3 && (c.flags & kGemmAKContiguous)))
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "an output shift needs the K-contiguous split-bf16 path");
    const Tile &T = kTiles[tile];
    *grid = dim3((c.M + T.bm - 1) / T.bm, (c.N + T.bn - 1) / T.bn, c.batch);
    if (precision == kF64) {
        if (c.typeA != kDouble || c.typeB != kDouble || c.typeC != kDouble || c.flags)
            return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "FP64 contraction takes plain double operands");
        *fn = tile == 0 ? f64_kernel<128, 128>() : (tile == 1 ? f64_kernel<64, 64>() : f64_kernel<32, 32>());
        return FORT_SCOPE_OK;
    }
    if (precision != kBf16x3) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown contraction precision");
    if (c.flags & kGemmAKContiguous) {
        const bool tout = (c.flags & kGemmTransposedC) != 0;
        const int combo = c.typeA * 27 + c.typeB * 9 + c.typeC * 3;
        constexpr int D = kDouble, F = kFloat, S = kSplit;
        if (tout) {
            switch (combo) {
            case F * 27 + D * 9 + F * 3: *fn = tn_by_tile<float, double, float, true>(tile); break;
            case F * 27 + F * 9 + F * 3: *fn = tn_by_tile<float, float, float, true>(tile); break;
            case S * 27 + D * 9 + S * 3: *fn = tn_by_tile<Split16, double, Split16, true>(tile); break;
            case S * 27 + S * 9 + F * 3: *fn = tn_by_tile<Split16, Split16, float, true>(tile); break;
            case S * 27 + S * 9 + S * 3: *fn = tn_by_tile<Split16, Split16, Split16, true>(tile); break;
            case S * 27 + D * 9 + F * 3: *fn = tn_by_tile<Split16, double, float, true>(tile); break;
            default: return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unsupported K-contiguous operand types");
            }
        } else {
            switch (combo) {
            case F * 27 + F * 9 + F * 3: *fn = tn_by_tile<float, float, float, false>(tile); break;
            case F * 27 + F * 9 + D * 3: *fn = tn_by_tile<float, float, double, false>(tile); break;
            case D * 27 + D * 9 + D * 3: *fn = tn_by_tile<double, double, double, false>(tile); break;
            case S * 27 + S * 9 + S * 3: *fn = tn_by_tile<Split16, Split16, Split16, false>(tile); break;
            case S * 27 + S * 9 + D * 3: *fn = tn_by_tile<Split16, Split16, double, false>(tile); break;
            case F * 27 + D * 9 + D * 3: *fn = tn_by_tile<float, double, double, false>(tile); break;
            case F * 27 + S * 9 + D * 3: *fn = tn_by_tile<float, Split16, double, false>(tile); break;
            default: return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unsupported K-contiguous operand types");
            }
        }
        return FORT_SCOPE_OK;
    }
    if (c.flags) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "a transposed C needs a K-contiguous A");
    if (c.typeA == kSplit || c.typeB == kSplit || c.typeC == kSplit)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "split operands need a K-contiguous A");
    // The operand type combinations the chain contraction and the spectral transforms use.
    const int combo = c.typeA * 4 + c.typeB * 2 + c.typeC;
    switch (combo) {
    case kDouble * 4 + kDouble * 2 + kDouble: *fn = bf16_by_tile<double, double, double>(tile); break;
    case kDouble * 4 + kDouble * 2 + kFloat: *fn = bf16_by_tile<double, double, float>(tile); break;
    case kFloat * 4 + kDouble * 2 + kFloat: *fn = bf16_by_tile<float, double, float>(tile); break;
    case kFloat * 4 + kFloat * 2 + kFloat: *fn = bf16_by_tile<float, float, float>(tile); break;
    case kFloat * 4 + kFloat * 2 + kDouble: *fn = bf16_by_tile<float, float, double>(tile); break;
    default: return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unsupported split-bf16 operand types");
    }
    return FORT_SCOPE_OK;
}

static int validate(const GemmCall &c)
{
    if (c.M < 0 || c.N < 0 || c.K < 0 || c.batch < 0)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "negative contraction extent");
    if (c.batch > 65535) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "batch above 65535");
    if (c.flags & ~(kGemmAKContiguous | kGemmTransposedC))
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown contraction layout flags");
    const long long amin = (c.flags & kGemmAKContiguous) ? c.K : c.M;
    const long long cmin = (c.flags & kGemmTransposedC) ? c.N : c.M;
    if ((amin > 0 && c.lda < amin) || (c.K > 0 && c.ldb < c.K) || (cmin > 0 && c.ldc < cmin))
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "leading dimension below the row count");
    if ((long long)(c.N + 31) / 32 > 65535)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "too many column tiles");
    return FORT_SCOPE_OK;
}

int gemm_launch(const GemmCall &c, int precision, hipStream_t stream, int tile)
{
    FTC_CHECK(validate(c));
    if (!c.M || !c.N || !c.batch) return FORT_SCOPE_OK;
    const void *fn = nullptr;
    dim3 grid;
    FTC_CHECK(select_kernel(c, precision, tile, &fn, &grid));
    GemmCall arg = c;
    void *params[1] = {&arg};
    if (hipLaunchKernel(fn, grid, dim3(kGemmThreads), params, 0, stream) != hipSuccess) {
        (void)hipGetLastError();
        return fort_scope_report_error(FORT_SCOPE_EXECUTION, "contraction launch failed");
    }
    return FORT_SCOPE_OK;
}

// Queues one contraction as a task of the context (no bindings: the caller binds its arrays in the
// tasks that produce or consume them, or passes them through bindings of its own).
static int gemm_enqueue(fort_scope_t ctx, const GemmCall &c, int precision, const std::vector<Binding3> &b)
{
    FTC_CHECK(validate(c));
    if (!c.M || !c.N || !c.batch) return FORT_SCOPE_OK;
    const void *fn = nullptr;
    dim3 grid;
    FTC_CHECK(select_kernel(c, precision, -1, &fn, &grid));
    return enqueue_kernel(ctx, 0x6e33000000000000ull | (uint64_t)precision, fn, grid, dim3(kGemmThreads), 0, &c,
                          sizeof c, b);
}

// ------------------------------------------------------------------------------------------------ chain
// out = A^T split: out[m*ldo + k] (bf16 hi) and out[lo + m*ldo + k] (bf16 lo) hold A(m, k) = a[m + k*lda]
// rounded to float and split as the contractions split it, so A becomes a K-contiguous split operand.
// 64 x 64 tiles through LDS: columns of A are read and rows of the output written as contiguous runs.
struct SplitTransposeArgs {
    const double *a;
    __bf16 *out;
    int m, k;
    long long lda, ldo, lo;
};

__global__ __launch_bounds__(256) void split_transpose_kernel(const SplitTransposeArgs p)
{
    __shared__ double tile[64][65];
    const int m0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int kk = ty + 4 * r, m = m0 + tx, k = k0 + kk;
        tile[kk][tx] = (m < p.m && k < p.k) ? p.a[(long long)m + (long long)k * p.lda] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int mm = ty + 4 * r, m = m0 + mm, k = k0 + tx;
        if (m < p.m && k < p.k) {
            const float f = (float)tile[tx][mm];
            const __bf16 h = (__bf16)f;
            const long long o = (long long)m * p.ldo + k;
            p.out[o] = h;
            p.out[p.lo + o] = (__bf16)(f - (float)h);
        }
    }
}

// out[i] (bf16 hi) and out[lo + i] (bf16 lo) hold a[i] rounded to float and split, for i < n: a
// column-major operand that is already K-contiguous, split once instead of in every tile that loads it.
struct SplitArgs {
    const double *a;
    __bf16 *out;
    long long n, lo;
};

__global__ __launch_bounds__(256) void split_planes_kernel(const SplitArgs p)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < p.n; i += (long long)gridDim.x * 256) {
        const float f = (float)p.a[i];
        const __bf16 h = (__bf16)f;
        p.out[i] = h;
        p.out[p.lo + i] = (__bf16)(f - (float)h);
    }
}

// ------------------------------------------------------------------------------------------------ probe
__global__ void mfma_probe_f64(double *out, int a_lane, int b_lane)
{
    const int l = threadIdx.x;
    const double a = a_lane < 0 ? 1.0 : (l == a_lane ? 1.0 : 0.0);
    const double b = b_lane < 0 ? 1.0 : (l == b_lane ? 1.0 : 0.0);
    f64x4 c = {0.0, 0.0, 0.0, 0.0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[l * 4 + r] = c[r];
}

__global__ void mfma_probe_bf16(float *out, int a_lane, int a_elem, int b_lane, int b_elem)
{
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) {
        a[e] = (__bf16)((a_lane < 0 || (l == a_lane && e == a_elem)) ? 1.0f : 0.0f);
        b[e] = (__bf16)((b_lane < 0 || (l == b_lane && e == b_elem)) ? 1.0f : 0.0f);
    }
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[l * 4 + r] = c[r];
}

}  // namespace ftc

using namespace ftc;

extern "C" {


int ftc_gemm_device(const ftc_gemm_desc *d, int precision, int tile, void *stream)
{
    if (!d) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "missing contraction descriptor");
    for (int t : {d->type_a, d->type_b, d->type_c})
        if (t != kDouble && t != kFloat) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown operand type");
    GemmCall c{d->a, d->b, d->c, d->type_a, d->type_b, d->type_c, d->m, d->n, d->k, d->batch, d->lda, d->ldb,
               d->ldc, d->stride_a, d->stride_b, d->stride_c, d->alpha, d->beta, d->flags};
    return gemm_launch(c, precision, static_cast<hipStream_t>(stream), tile);
}

int ftc_gemm_tile_count(void) { return kTileCount; }

// One MFMA with one-hot (or all-ones, lane < 0) operands; writes the 64 x 4 accumulator lanes.
// kind 0: f64 16x16x4 (elem ignored), kind 1: bf16 16x16x32 (output converted to double).
int ftc_mfma_layout_probe(int kind, int a_lane, int a_elem, int b_lane, int b_elem, double *out256)
{
    if (!out256 || a_lane > 63 || b_lane > 63 || a_elem < 0 || a_elem > 7 || b_elem < 0 || b_elem > 7)
        return FORT_SCOPE_ARGUMENT;
    void *d = nullptr;
    if (hipMalloc(&d, 256 * sizeof(double)) != hipSuccess) return FORT_SCOPE_RESOURCE;
    int status = FORT_SCOPE_OK;
    if (kind == 0) {
        hipLaunchKernelGGL(mfma_probe_f64, dim3(1), dim3(64), 0, 0, static_cast<double *>(d), a_lane, b_lane);
        if (hipMemcpy(out256, d, 256 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) status = FORT_SCOPE_EXECUTION;
    } else {
        hipLaunchKernelGGL(mfma_probe_bf16, dim3(1), dim3(64), 0, 0, static_cast<float *>(d), a_lane, a_elem, b_lane,
                           b_elem);
        float tmp[256];
        if (hipMemcpy(tmp, d, 256 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) status = FORT_SCOPE_EXECUTION;
        for (int e = 0; e < 256; ++e) out256[e] = tmp[e];
    }
    (void)hipFree(d);
    return status;
}

// D = A * B * C for n x n column-major FP64 arrays (the chain of PolyBench 2mm/3mm contractions,
// compiler/tests/test_polybench.py:27-41): T = A*B into a scratch lease, then D = T*C.
// precision 0: FP64 MFMA; 1: split-bf16 products with FP32 accumulation.
int cpp_fort_shared_chain_contraction(fort_scope_t ctx, int mode, fort_buffer_t d, fort_buffer_t a,
                                      fort_buffer_t b, fort_buffer_t c, const int *pn, const int *pprecision)
{
    FTC_CHECK(entry_preamble(mode));
    if (!pn || !pprecision) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    const int n = *pn, precision = *pprecision;
    if (precision != kF64 && precision != kBf16x3)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown contraction precision");
    for (fort_buffer_t h : {d, a, b, c}) {
        fort_scope_layout l{};
        FTC_CHECK(fort_scope_layout_get(ctx, h, &l));
        if (l.rank != 2 || l.type != FORT_SCOPE_REAL64 || l.element_bytes != 8)
            return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "entry layout/type mismatch");
        if (l.extents[0] != (std::size_t)n || l.extents[1] != (std::size_t)n)
            return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "contraction arrays must be n x n");
    }
    FTC_CHECK(distinct_outputs({d}, {a, b, c}));
    if (n <= 0) return FORT_SCOPE_OK;
    FTC_CHECK(fort_scope_forget_definition(ctx, d));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {d, a, b, c}, dev));
    // FP64: T = A*B into an FP64 scratch array, then D = T*C.
    // Split-bf16: every operand K-contiguous and split once. A is transposed and split (At), B and C
    // (column-major, so already K-contiguous) are split in place order (Bs, Cs), and T = A*B is written
    // transposed in FP32 (row-major: K-contiguous as the A operand of the second product). Scratch: At, Bs,
    // Cs and T, n*n*4 bytes each.
    const std::size_t tbytes = (std::size_t)n * n * (precision == kF64 ? 8 : 16);
    fort_scope_scratch_lease_v1 lease{};
    FTC_CHECK(fort_scope_scratch_acquire_v1(ctx, tbytes, &lease));
    const std::size_t N = n;
    const Sec3 all2{{0, 0, 0}, {N, N, 1}};
    const long long nn = (long long)n * n;
    int status = FORT_SCOPE_OK;
    // Binding sections are rank 3 in the helper; a rank-2 array uses the first two bounds.
    if (precision == kF64) {
        GemmCall first{dev[1], dev[2], lease.device, kDouble, kDouble, kDouble, n, n, n, 1, n, n, n, 0, 0, 0, 1.0, 0.0};
        GemmCall second{lease.device, dev[3], dev[0], kDouble, kDouble, kDouble, n, n, n, 1, n, n, n, 0, 0, 0, 1.0, 0.0};
        status = gemm_enqueue(ctx, first, precision, {{a, {all2}, {}, {}}, {b, {all2}, {}, {}}});
        if (status == FORT_SCOPE_OK)
            status = gemm_enqueue(ctx, second, precision, {{c, {all2}, {}, {}}, {d, {}, {}, {all2}}});
    } else {
        __bf16 *at = static_cast<__bf16 *>(lease.device), *bs = at + 2 * nn, *cs = bs + 2 * nn;
        float *t = reinterpret_cast<float *>(cs + 2 * nn);
        const SplitTransposeArgs st{static_cast<const double *>(dev[1]), at, n, n, n, n, nn};
        status = enqueue_kernel(ctx, 0x5e71000000000000ull, reinterpret_cast<const void *>(&split_transpose_kernel),
                                dim3((n + 63) / 64, (n + 63) / 64), dim3(256), 0, &st, sizeof st,
                                {{a, {all2}, {}, {}}});
        const unsigned sgrid = (unsigned)std::min<long long>((nn + 255) / 256, 4096);
        const SplitArgs sb{static_cast<const double *>(dev[2]), bs, nn, nn}, sc{static_cast<const double *>(dev[3]), cs, nn, nn};
        if (status == FORT_SCOPE_OK)
            status = enqueue_kernel(ctx, 0x5e72000000000000ull, reinterpret_cast<const void *>(&split_planes_kernel),
                                    dim3(sgrid), dim3(256), 0, &sb, sizeof sb, {{b, {all2}, {}, {}}});
        if (status == FORT_SCOPE_OK)
            status = enqueue_kernel(ctx, 0x5e72000000000000ull, reinterpret_cast<const void *>(&split_planes_kernel),
                                    dim3(sgrid), dim3(256), 0, &sc, sizeof sc, {{c, {all2}, {}, {}}});
        const int ak = kGemmAKContiguous, at_flags = kGemmAKContiguous | kGemmTransposedC;
        GemmCall first{at, bs, t, kSplit, kSplit, kFloat, n, n, n, 1, n, n, n, 0, 0, 0, 1.0, 0.0, at_flags, nn, nn, 0};
        GemmCall second{t, cs, dev[0], kFloat, kSplit, kDouble, n, n, n, 1, n, n, n, 0, 0, 0, 1.0, 0.0, ak, 0, nn, 0};
        if (status == FORT_SCOPE_OK) status = gemm_enqueue(ctx, first, precision, {});
        if (status == FORT_SCOPE_OK) status = gemm_enqueue(ctx, second, precision, {{d, {}, {}, {all2}}});
    }
    int released = fort_scope_scratch_release_v1(ctx, lease.token);
    return status ? status : released;
}

}  // extern "C"
