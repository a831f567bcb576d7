// MFMA contractions for gfx950 (see gemm.hpp for the operation and precisions), the dense chain
// contraction entry D = A*B*C, and an in-bounds probe of the MFMA lane layouts the kernels assume.
//
// Lane layouts used (wave64, l = lane):
//   v_mfma_f64_16x16x4f64   A(16x4):  lane l holds A[l%16][l/16];  B(4x16): B[l/16][l%16];
//                           D(16x16): lane l, element r holds D[l/16 + 4*r][l%16]
//   v_mfma_f32_16x16x32_bf16 A(16x32): lane l holds A[l%16][8*(l/16) .. 8*(l/16)+7];
//                           B(32x16): B[8*(l/16) .. +7][l%16];
//                           D(16x16, FP32): lane l, element r holds D[4*(l/16) + r][l%16]
// tests/test_dense_gpu.py checks both against the probe before trusting any product.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ftc_dense.h"
#include "entry.hpp"
#include "gemm.hpp"

namespace ftc {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kGemmThreads = 256;

// ------------------------------------------------------------------------------------------------ FP64
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(kGemmThreads) void gemm_f64_kernel(const GemmCall g)
{
    constexpr int BK = 16;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TM = WTM / 16, TN = WTN / 16;
    static_assert(WM * WN == kGemmThreads / 64, "four waves");
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const long long b = blockIdx.z;
    const double *A = static_cast<const double *>(g.A) + b * g.sa;
    const double *B = static_cast<const double *>(g.B) + b * g.sb;
    f64x4 acc[TM][TN];
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        double ra[BM * BK / kGemmThreads], rb[BN * BK / kGemmThreads];
#pragma unroll
        for (int e = 0; e < BM * BK / kGemmThreads; ++e) {
            const int idx = t + e * kGemmThreads, m = idx % BM, k = idx / BM;
            const bool ok = m0 + m < g.M && k0 + k < g.K;
            ra[e] = ok ? A[(long long)(m0 + m) + (long long)(k0 + k) * g.lda] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < BN * BK / kGemmThreads; ++e) {
            const int idx = t + e * kGemmThreads, k = idx % BK, n = idx / BK;
            const bool ok = n0 + n < g.N && k0 + k < g.K;
            rb[e] = ok ? B[(long long)(k0 + k) + (long long)(n0 + n) * g.ldb] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < BM * BK / kGemmThreads; ++e) {
            const int idx = t + e * kGemmThreads;
            As[idx / BM][idx % BM] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < BN * BK / kGemmThreads; ++e) {
            const int idx = t + e * kGemmThreads;
            Bs[idx % BK][idx / BK] = rb[e];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double af[TM], bfr[TN];
#pragma unroll
            for (int x = 0; x < TM; ++x) af[x] = As[kk + lane / 16][wm * WTM + x * 16 + lane % 16];
#pragma unroll
            for (int y = 0; y < TN; ++y) bfr[y] = Bs[kk + lane / 16][wn * WTN + y * 16 + lane % 16];
#pragma unroll
            for (int x = 0; x < TM; ++x)
#pragma unroll
                for (int y = 0; y < TN; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[x], bfr[y], acc[x][y], 0, 0, 0);
        }
    }
    double *C = static_cast<double *>(g.C) + b * g.sc;
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * WTM + x * 16 + lane / 16 + 4 * r;
                const int col = n0 + wn * WTN + y * 16 + lane % 16;
                if (row < g.M && col < g.N) {
                    const long long ci = (long long)row + (long long)col * g.ldc;
                    double v = g.alpha * acc[x][y][r];
                    if (g.beta != 0.0) v = v + g.beta * C[ci];
                    C[ci] = v;
                }
            }
}

// ------------------------------------------------------------------------------------------------ bf16x3
template <class TA, class TB, class TO, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(kGemmThreads) void gemm_bf16x3_kernel(const GemmCall g)
{
    constexpr int BK = 32, PAD = 8;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TM = WTM / 16, TN = WTN / 16;
    static_assert(WM * WN == kGemmThreads / 64, "four waves");
    __shared__ __attribute__((aligned(16))) __bf16 Ah[BM][BK + PAD];
    __shared__ __attribute__((aligned(16))) __bf16 Al[BM][BK + PAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[BN][BK + PAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[BN][BK + PAD];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const long long b = blockIdx.z;
    const TA *A = static_cast<const TA *>(g.A) + b * g.sa;
    const TB *B = static_cast<const TB *>(g.B) + b * g.sb;
    f32x4 acc[TM][TN];
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) acc[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        float ra[BM * BK / kGemmThreads], rb[BN * BK / kGemmThreads];
#pragma unroll
        for (int e = 0; e < BM * BK / kGemmThreads; ++e) {
            const int idx = t + e * kGemmThreads, m = idx % BM, k = idx / BM;
            const bool ok = m0 + m < g.M && k0 + k < g.K;
            ra[e] = ok ? (float)A[(long long)(m0 + m) + (long long)(k0 + k) * g.lda] : 0.f;
        }
#pragma unroll
        for (int e = 0; e < BN * BK / kGemmThreads; ++e) {
            const int idx = t + e * kGemmThreads, k = idx % BK, n = idx / BK;
            const bool ok = n0 + n < g.N && k0 + k < g.K;
            rb[e] = ok ? (float)B[(long long)(k0 + k) + (long long)(n0 + n) * g.ldb] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < BM * BK / kGemmThreads; ++e) {
            const int idx = t + e * kGemmThreads, m = idx % BM, k = idx / BM;
            const __bf16 hi = (__bf16)ra[e];
            Ah[m][k] = hi;
            Al[m][k] = (__bf16)(ra[e] - (float)hi);
        }
#pragma unroll
        for (int e = 0; e < BN * BK / kGemmThreads; ++e) {
            const int idx = t + e * kGemmThreads, k = idx % BK, n = idx / BK;
            const __bf16 hi = (__bf16)rb[e];
            Bh[n][k] = hi;
            Bl[n][k] = (__bf16)(rb[e] - (float)hi);
        }
        __syncthreads();
        bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
        for (int x = 0; x < TM; ++x) {
            const int row = wm * WTM + x * 16 + lane % 16;
            ah[x] = *reinterpret_cast<const bf16x8 *>(&Ah[row][8 * (lane / 16)]);
            al[x] = *reinterpret_cast<const bf16x8 *>(&Al[row][8 * (lane / 16)]);
        }
#pragma unroll
        for (int y = 0; y < TN; ++y) {
            const int col = wn * WTN + y * 16 + lane % 16;
            bh[y] = *reinterpret_cast<const bf16x8 *>(&Bh[col][8 * (lane / 16)]);
            bl[y] = *reinterpret_cast<const bf16x8 *>(&Bl[col][8 * (lane / 16)]);
        }
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
            for (int y = 0; y < TN; ++y) {
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
            }
    }
    TO *C = static_cast<TO *>(g.C) + b * g.sc;
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * WTM + x * 16 + 4 * (lane / 16) + r;
                const int col = n0 + wn * WTN + y * 16 + lane % 16;
                if (row < g.M && col < g.N) {
                    const long long ci = (long long)row + (long long)col * g.ldc;
                    double v = g.alpha * (double)acc[x][y][r];
                    if (g.beta != 0.0) v = v + g.beta * (double)C[ci];
                    C[ci] = (TO)v;
                }
            }
}

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
    const Tile &T = kTiles[tile];
    *grid = dim3((c.M + T.bm - 1) / T.bm, (c.N + T.bn - 1) / T.bn, c.batch);
    if (precision == kF64) {
        if (c.typeA != kDouble || c.typeB != kDouble || c.typeC != kDouble)
            return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "FP64 contraction takes double operands");
        *fn = tile == 0 ? f64_kernel<128, 128>() : (tile == 1 ? f64_kernel<64, 64>() : f64_kernel<32, 32>());
        return FORT_SCOPE_OK;
    }
    if (precision != kBf16x3) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown contraction precision");
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
    if ((c.M > 0 && c.lda < c.M) || (c.K > 0 && c.ldb < c.K) || (c.M > 0 && c.ldc < c.M))
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
    GemmCall c{d->a, d->b, d->c, d->type_a, d->type_b, d->type_c, d->m, d->n, d->k, d->batch, d->lda, d->ldb,
               d->ldc, d->stride_a, d->stride_b, d->stride_c, d->alpha, d->beta};
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
    // T = A*B stays FP64 in both precisions: its traffic is negligible next to the 4n^3 flops.
    const std::size_t tbytes = (std::size_t)n * n * 8;
    fort_scope_scratch_lease_v1 lease{};
    FTC_CHECK(fort_scope_scratch_acquire_v1(ctx, tbytes, &lease));
    const std::size_t N = n;
    const Sec3 all2{{0, 0, 0}, {N, N, 1}};
    GemmCall first{dev[1], dev[2], lease.device, kDouble, kDouble, kDouble, n, n, n, 1, n, n, n, 0, 0, 0, 1.0, 0.0};
    GemmCall second{lease.device, dev[3], dev[0], kDouble, kDouble, kDouble, n, n, n, 1, n, n, n, 0, 0, 0, 1.0, 0.0};
    // Binding sections are rank 3 in the helper; a rank-2 array uses the first two bounds.
    int status = gemm_enqueue(ctx, first, precision, {{a, {all2}, {}, {}}, {b, {all2}, {}, {}}});
    if (status == FORT_SCOPE_OK)
        status = gemm_enqueue(ctx, second, precision, {{c, {all2}, {}, {}}, {d, {}, {}, {all2}}});
    int released = fort_scope_scratch_release_v1(ctx, lease.token);
    return status ? status : released;
}

}  // extern "C"
