// Device code of the MFMA contractions (gemm.hpp describes the operation): per-tile bodies, used by the
// one-call kernels in gemm.hip and by the persistent grouped launches of the ensemble.
//
// Lane layouts used (wave64, l = lane):
//   v_mfma_f64_16x16x4f64   A(16x4):  lane l holds A[l%16][l/16];  B(4x16): B[l/16][l%16];
//                           D(16x16): lane l, element r holds D[l/16 + 4*r][l%16]
//   v_mfma_f32_16x16x32_bf16 A(16x32): lane l holds A[l%16][8*(l/16) .. 8*(l/16)+7];
//                           B(32x16): B[8*(l/16) .. +7][l%16];
//                           D(16x16, FP32): lane l, element r holds D[4*(l/16) + r][l%16]
// tests/test_dense_gpu.py checks both against the probe before trusting any product.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "gemm.hpp"
#include "xcd.hpp"

namespace ftc {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kGemmThreads = 256;

// ------------------------------------------------------------------------------------------------ FP64
template <int BM, int BN>
struct F64Lds {
    double As[16][BM + 1];
    double Bs[16][BN + 1];
};

// One BM x BN tile (mt, nt) of batch entry b of an FP64 contraction. The global loads of K-slice s+1 are
// issued before the MFMAs of slice s, so their latency hides behind the products. The MFMA operands are
// swapped (the accumulators hold C^T tiles: lane l, element r is C(l%16, l/16 + 4r)), so sixteen
// consecutive lanes store sixteen consecutive rows of the column-major C.
template <int BM, int BN, int WM, int WN>
__device__ __forceinline__ void gemm_f64_tile(const GemmCall &g, F64Lds<BM, BN> &L, int mt, int nt, long long b)
{
    constexpr int BK = 16;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TM = WTM / 16, TN = WTN / 16;
    constexpr int NA = BM * BK / kGemmThreads, NB = BN * BK / kGemmThreads;
    static_assert(WM * WN == kGemmThreads / 64, "four waves");
    auto &As = L.As;
    auto &Bs = L.Bs;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int m0 = mt * BM, n0 = nt * BN;
    const double *A = static_cast<const double *>(g.A) + b * g.sa;
    const double *B = static_cast<const double *>(g.B) + b * g.sb;
    f64x4 acc[TM][TN];
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
    double ra[NA], rb[NB];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < NA; ++e) {
            const int idx = t + e * kGemmThreads, m = idx % BM, k = idx / BM;
            const bool ok = m0 + m < g.M && k0 + k < g.K;
            ra[e] = ok ? A[(long long)(m0 + m) + (long long)(k0 + k) * g.lda] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            const int idx = t + e * kGemmThreads, k = idx % BK, n = idx / BK;
            const bool ok = n0 + n < g.N && k0 + k < g.K;
            rb[e] = ok ? B[(long long)(k0 + k) + (long long)(n0 + n) * g.ldb] : 0.0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NA; ++e) {
            const int idx = t + e * kGemmThreads;
            As[idx / BM][idx % BM] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            const int idx = t + e * kGemmThreads;
            Bs[idx % BK][idx / BK] = rb[e];
        }
        __syncthreads();
        if (k0 + BK < g.K) load(k0 + BK);
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
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bfr[y], af[x], acc[x][y], 0, 0, 0);
        }
    }
    double *C = static_cast<double *>(g.C) + b * g.sc;
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * WTM + x * 16 + lane % 16;
                const int col = n0 + wn * WTN + y * 16 + lane / 16 + 4 * r;
                if (row < g.M && col < g.N) {
                    const long long ci = (long long)row + (long long)col * g.ldc;
                    double v = g.alpha * acc[x][y][r];
                    if (g.beta != 0.0) v = v + g.beta * C[ci];
                    C[ci] = v;
                }
            }
}

// The same contraction with two LDS stages of BK elements of K: slice s+1 is stored into the other stage
// while slice s is multiplied, so each slice costs one barrier, and the global loads of slice s+2 are in
// flight meanwhile. Summation order over K is that of gemm_f64_tile (k ascending, four per MFMA).
template <int BM, int BN, int BK>
struct F64DbLds {
    double As[2][BK][BM + 1];
    double Bs[2][BK][BN + 1];
};

template <int BM, int BN, int WM, int WN, int BK>
__device__ __forceinline__ void gemm_f64_db_tile(const GemmCall &g, F64DbLds<BM, BN, BK> &L, int mt, int nt,
                                                 long long b)
{
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TM = WTM / 16, TN = WTN / 16;
    constexpr int NA = BM * BK / kGemmThreads, NB = BN * BK / kGemmThreads;
    static_assert(WM * WN == kGemmThreads / 64, "four waves");
    static_assert(BK % 4 == 0 && NA * kGemmThreads == BM * BK && NB * kGemmThreads == BN * BK, "whole stages");
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int m0 = mt * BM, n0 = nt * BN;
    const double *A = static_cast<const double *>(g.A) + b * g.sa;
    const double *B = static_cast<const double *>(g.B) + b * g.sb;
    f64x4 acc[TM][TN];
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
    double ra[NA], rb[NB];
    // A tile whose rows, columns and K-slices all lie inside the operands loads without bounds checks.
    const bool whole = m0 + BM <= g.M && n0 + BN <= g.N && g.K % BK == 0;
    auto load = [&](int k0) {
        if (whole) {
#pragma unroll
            for (int e = 0; e < NA; ++e) {
                const int idx = t + e * kGemmThreads, m = idx % BM, k = idx / BM;
                ra[e] = A[(long long)(m0 + m) + (long long)(k0 + k) * g.lda];
            }
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                const int idx = t + e * kGemmThreads, k = idx % BK, n = idx / BK;
                rb[e] = B[(long long)(k0 + k) + (long long)(n0 + n) * g.ldb];
            }
            return;
        }
#pragma unroll
        for (int e = 0; e < NA; ++e) {
            const int idx = t + e * kGemmThreads, m = idx % BM, k = idx / BM;
            const bool ok = m0 + m < g.M && k0 + k < g.K;
            ra[e] = ok ? A[(long long)(m0 + m) + (long long)(k0 + k) * g.lda] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            const int idx = t + e * kGemmThreads, k = idx % BK, n = idx / BK;
            const bool ok = n0 + n < g.N && k0 + k < g.K;
            rb[e] = ok ? B[(long long)(k0 + k) + (long long)(n0 + n) * g.ldb] : 0.0;
        }
    };
    auto put = [&](int stage) {
#pragma unroll
        for (int e = 0; e < NA; ++e) {
            const int idx = t + e * kGemmThreads;
            L.As[stage][idx / BM][idx % BM] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            const int idx = t + e * kGemmThreads;
            L.Bs[stage][idx % BK][idx / BK] = rb[e];
        }
    };
    // The operand fragments of MFMA step kk+4 are read from LDS while the products of step kk run.
    auto multiply = [&](int stage) {
        double af[2][TM], bfr[2][TN];
        auto fragments = [&](int buf, int kk) {
#pragma unroll
            for (int x = 0; x < TM; ++x) af[buf][x] = L.As[stage][kk + lane / 16][wm * WTM + x * 16 + lane % 16];
#pragma unroll
            for (int y = 0; y < TN; ++y) bfr[buf][y] = L.Bs[stage][kk + lane / 16][wn * WTN + y * 16 + lane % 16];
        };
        fragments(0, 0);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int cur = (kk / 4) & 1;
            if (kk + 4 < BK) fragments(cur ^ 1, kk + 4);
#pragma unroll
            for (int x = 0; x < TM; ++x)
#pragma unroll
                for (int y = 0; y < TN; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bfr[cur][y], af[cur][x], acc[x][y], 0, 0, 0);
        }
    };
    load(0);
    put(0);
    if (BK < g.K) load(BK);
    __syncthreads();
    int stage = 0, k0 = 0;
    // Every slice but the last: multiply, store the next slice into the other stage (its readers passed
    // the previous barrier), start the loads of the slice after. The last slice is peeled off, so the
    // loop body has no store branch.
    for (; k0 + BK < g.K; k0 += BK) {
        multiply(stage);
        put(stage ^ 1);
        if (k0 + 2 * BK < g.K) load(k0 + 2 * BK);
        __syncthreads();
        stage ^= 1;
    }
    multiply(stage);
    double *C = static_cast<double *>(g.C) + b * g.sc;
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * WTM + x * 16 + lane % 16;
                const int col = n0 + wn * WTN + y * 16 + lane / 16 + 4 * r;
                if (row < g.M && col < g.N) {
                    const long long ci = (long long)row + (long long)col * g.ldc;
                    double v = g.alpha * acc[x][y][r];
                    if (g.beta != 0.0) v = v + g.beta * C[ci];
                    C[ci] = v;
                }
            }
}

template <int BM, int BN, int WM, int WN, int BK>
__global__ __launch_bounds__(kGemmThreads) void gemm_f64_db_kernel(const GemmCall g)
{
    __shared__ F64DbLds<BM, BN, BK> L;
    int mt, nt;
    long long b;
    gemm_block_tile(mt, nt, b);
    gemm_f64_db_tile<BM, BN, WM, WN, BK>(g, L, mt, nt, b);
}

// Tiles below 128 x 128 declare four waves per SIMD: with a larger register budget the compiler keeps their
// accumulators in VGPRs and copies them through AGPRs around every group of MFMAs.
constexpr int gemm_small_waves(int bm) { return bm >= 128 ? 1 : 4; }

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(kGemmThreads) __attribute__((amdgpu_waves_per_eu(gemm_small_waves(BM)))) void
gemm_f64_kernel(const GemmCall g)
{
    __shared__ F64Lds<BM, BN> L;
    int mt, nt;
    long long b;
    gemm_block_tile(mt, nt, b);
    gemm_f64_tile<BM, BN, WM, WN>(g, L, mt, nt, b);
}

// ------------------------------------------------------------------------------------------------ bf16x3
template <int BM, int BN>
struct Bf16Lds {
    __attribute__((aligned(16))) __bf16 Ah[BM][40];
    __attribute__((aligned(16))) __bf16 Al[BM][40];
    __attribute__((aligned(16))) __bf16 Bh[BN][40];
    __attribute__((aligned(16))) __bf16 Bl[BN][40];
};

// One BM x BN tile (mt, nt) of batch entry b of a split-bf16 contraction.
template <class TA, class TB, class TO, int BM, int BN, int WM, int WN>
__device__ __forceinline__ void gemm_bf16x3_tile(const GemmCall &g, Bf16Lds<BM, BN> &L, int mt, int nt, long long b)
{
    constexpr int BK = 32;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TM = WTM / 16, TN = WTN / 16;
    static_assert(WM * WN == kGemmThreads / 64, "four waves");
    auto &Ah = L.Ah;
    auto &Al = L.Al;
    auto &Bh = L.Bh;
    auto &Bl = L.Bl;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int m0 = mt * BM, n0 = nt * BN;
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

template <class TA, class TB, class TO, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(kGemmThreads) __attribute__((amdgpu_waves_per_eu(gemm_small_waves(BM)))) void
gemm_bf16x3_kernel(const GemmCall g)
{
    __shared__ Bf16Lds<BM, BN> L;
    int mt, nt;
    long long b;
    gemm_block_tile(mt, nt, b);
    gemm_bf16x3_tile<TA, TB, TO, BM, BN, WM, WN>(g, L, mt, nt, b);
}

// ------------------------------------------------------------------------------------------------ bf16x3, K-contiguous
// Split-bf16 tile with both operands K-contiguous: A(m, k) = A[m*lda + k], B(k, n) = B[k + n*ldb]. Four
// consecutive threads read one 32-element K run of a row (one or two cache lines), split it in registers
// and store 16-byte hi and lo vectors straight into the MFMA fragment layout. TOUT writes C transposed.
template <class T>
__device__ __forceinline__ void split8(const T (&v)[8], bf16x8 &hi, bf16x8 &lo)
{
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float f = (float)v[q];
        const __bf16 h = (__bf16)f;
        hi[q] = h;
        lo[q] = (__bf16)(f - (float)h);
    }
}

// 16-byte vectors that only assume the alignment of their elements: gfx950 serves a global dwordx4 load
// at any dword-aligned address, so a run that starts at an odd element (an interior that follows one ghost
// cell) still moves in 16-byte pieces.
typedef float f32x4_elem_aligned __attribute__((ext_vector_type(4), aligned(4)));
typedef double f64x2_elem_aligned __attribute__((ext_vector_type(2), aligned(8)));

// Eight consecutive elements src[0..valid) (zeros past `valid` or when !ok); 16-byte vector reads when the
// run is whole.
template <class T>
__device__ __forceinline__ void load8(const T *src, bool ok, int valid, T (&v)[8])
{
    if (ok && valid >= 8) {
        if constexpr (sizeof(T) == 4) {
            const f32x4_elem_aligned a = reinterpret_cast<const f32x4_elem_aligned *>(src)[0];
            const f32x4_elem_aligned b = reinterpret_cast<const f32x4_elem_aligned *>(src)[1];
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const f64x2_elem_aligned a = reinterpret_cast<const f64x2_elem_aligned *>(src)[q];
                v[2 * q] = a.x;
                v[2 * q + 1] = a.y;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = (ok && q < valid) ? src[q] : (T)0;
    }
}

// A kSplit operand: bf16 hi and lo planes (lo `lo` elements after hi).
struct Split16 {};

// Eight consecutive K elements of one operand row as staged in registers between the global load and the
// LDS store: raw values (split on the way into LDS) or, for Split16, the hi and lo vectors themselves.
template <class T>
struct Raw8 {
    T v[8];
};
template <>
struct Raw8<Split16> {
    bf16x8 h, l;
};

template <class T>
__device__ __forceinline__ void load_raw8(const void *base, long long idx, long long lo, bool ok, int valid, Raw8<T> &r)
{
    (void)lo;
    load8(static_cast<const T *>(base) + idx, ok, valid, r.v);
}

template <>
__device__ __forceinline__ void load_raw8<Split16>(const void *base, long long idx, long long lo, bool ok, int valid,
                                                   Raw8<Split16> &r)
{
    const __bf16 *h = static_cast<const __bf16 *>(base) + idx, *l = h + lo;
    if (ok && valid >= 8 && ((reinterpret_cast<uintptr_t>(h) | reinterpret_cast<uintptr_t>(l)) & 15) == 0) {
        r.h = *reinterpret_cast<const bf16x8 *>(h);
        r.l = *reinterpret_cast<const bf16x8 *>(l);
    } else {
        const __bf16 z = (__bf16)0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            r.h[q] = (ok && q < valid) ? h[q] : z;
            r.l[q] = (ok && q < valid) ? l[q] : z;
        }
    }
}

template <class T>
__device__ __forceinline__ void store_raw8(const Raw8<T> &r, __bf16 *hi, __bf16 *lo)
{
    bf16x8 h, l;
    split8(r.v, h, l);
    *reinterpret_cast<bf16x8 *>(hi) = h;
    *reinterpret_cast<bf16x8 *>(lo) = l;
}

template <>
__device__ __forceinline__ void store_raw8<Split16>(const Raw8<Split16> &r, __bf16 *hi, __bf16 *lo)
{
    *reinterpret_cast<bf16x8 *>(hi) = r.h;
    *reinterpret_cast<bf16x8 *>(lo) = r.l;
}

// Output element access: plain types, or Split16 (the value rounded to float, then split as split8 does).
template <class TO>
__device__ __forceinline__ double load_out(const void *C, long long ci, long long lo)
{
    (void)lo;
    return (double)static_cast<const TO *>(C)[ci];
}
template <>
__device__ __forceinline__ double load_out<Split16>(const void *C, long long ci, long long lo)
{
    const __bf16 *c = static_cast<const __bf16 *>(C);
    return (double)((float)c[ci] + (float)c[ci + lo]);
}
template <class TO>
__device__ __forceinline__ void store_out(void *C, long long ci, long long lo, double v)
{
    (void)lo;
    static_cast<TO *>(C)[ci] = (TO)v;
}
template <>
__device__ __forceinline__ void store_out<Split16>(void *C, long long ci, long long lo, double v)
{
    __bf16 *c = static_cast<__bf16 *>(C);
    const float f = (float)v;
    const __bf16 h = (__bf16)f;
    c[ci] = h;
    c[ci + lo] = (__bf16)(f - (float)h);
}

// Four adjacent output elements C[ci .. ci+3]: 16-byte vector accesses for FP32 and FP64 outputs (which
// need only element alignment), two 8-byte plane accesses for Split16 when ci is even.
template <class TO>
__device__ __forceinline__ void load_out4(const void *C, long long ci, long long lo, double (&v)[4])
{
    if constexpr (std::is_same<TO, double>::value) {
        const f64x2_elem_aligned *p = reinterpret_cast<const f64x2_elem_aligned *>(static_cast<const double *>(C) + ci);
        const f64x2_elem_aligned a = p[0], b = p[1];
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    } else if constexpr (std::is_same<TO, float>::value) {
        const f32x4_elem_aligned a = *reinterpret_cast<const f32x4_elem_aligned *>(static_cast<const float *>(C) + ci);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = load_out<TO>(C, ci + r, lo);
    }
}

typedef __bf16 bf16x4_dword_aligned __attribute__((ext_vector_type(4), aligned(4)));

template <class TO>
__device__ __forceinline__ void store_out4(void *C, long long ci, long long lo, const double (&v)[4])
{
    if constexpr (std::is_same<TO, double>::value) {
        f64x2_elem_aligned *p = reinterpret_cast<f64x2_elem_aligned *>(static_cast<double *>(C) + ci);
        p[0] = f64x2_elem_aligned{v[0], v[1]};
        p[1] = f64x2_elem_aligned{v[2], v[3]};
    } else if constexpr (std::is_same<TO, float>::value) {
        *reinterpret_cast<f32x4_elem_aligned *>(static_cast<float *>(C) + ci) =
            f32x4_elem_aligned{(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
    } else if constexpr (std::is_same<TO, Split16>::value) {
        if (((ci | lo) & 1) == 0) {
            bf16x4_dword_aligned h, l;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float f = (float)v[r];
                h[r] = (__bf16)f;
                l[r] = (__bf16)(f - (float)h[r]);
            }
            __bf16 *c = static_cast<__bf16 *>(C);
            *reinterpret_cast<bf16x4_dword_aligned *>(c + ci) = h;
            *reinterpret_cast<bf16x4_dword_aligned *>(c + ci + lo) = l;
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) store_out<TO>(C, ci + r, lo, v[r]);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) store_out<TO>(C, ci + r, lo, v[r]);
    }
}

template <class T>
__device__ __forceinline__ const void *offset_elems(const void *p, long long e)
{
    using E = typename std::conditional<std::is_same<T, Split16>::value, __bf16, T>::type;
    return static_cast<const E *>(p) + e;
}

// LDS of the K-contiguous tile: BK elements of K per row in 16-byte groups of 8, rows padded by 8 elements
// so that the sixteen rows one ds_read_b128 touches fall on disjoint banks. (An unpadded XOR-swizzled row
// of 64 left the bank-conflict counter unchanged and made the FP64-input transform 40 % slower.)
template <int BK>
struct TnRow {
    static constexpr int kWidth = BK + 8;
    __device__ static __forceinline__ int at(int, int group) { return 8 * group; }
};

template <int BM, int BN, int BK>
struct Bf16TnLds {
    __attribute__((aligned(16))) __bf16 Ah[BM][TnRow<BK>::kWidth];
    __attribute__((aligned(16))) __bf16 Al[BM][TnRow<BK>::kWidth];
    __attribute__((aligned(16))) __bf16 Bh[BN][TnRow<BK>::kWidth];
    __attribute__((aligned(16))) __bf16 Bl[BN][TnRow<BK>::kWidth];
};

// The epilogue of a K-contiguous split-bf16 tile: C = alpha*acc (+ beta*C) (- *shift), four adjacent
// elements per lane (see the operand order in gemm_bf16x3_tn_tile), vector accesses for whole runs.
template <class TO, int TM, int TN, int WTM, int WTN, bool TOUT>
__device__ __forceinline__ void tn_store_tile(const GemmCall &g, const f32x4 (&acc)[TM][TN], int m0, int n0, int wm,
                                              int wn, int lane, long long b)
{
    void *C = const_cast<void *>(offset_elems<TO>(g.C, b * g.sc));
    const bool shifted = g.shift != nullptr;
    const double shift = shifted ? *g.shift : 0.0;
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) {
            // (rr, cc) of element r = 0; element r is r rows (column-major) or r columns (TOUT) further on.
            const int rr = m0 + wm * WTM + x * 16 + (TOUT ? lane % 16 : 4 * (lane / 16));
            const int cc = n0 + wn * WTN + y * 16 + (TOUT ? 4 * (lane / 16) : lane % 16);
            const long long ci = TOUT ? (long long)rr * g.ldc + cc : (long long)rr + (long long)cc * g.ldc;
            const bool whole = TOUT ? rr < g.M && cc + 3 < g.N : rr + 3 < g.M && cc < g.N;
            if (whole) {
                double v[4];
                if (g.beta != 0.0) load_out4<TO>(C, ci, g.lo_c, v);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    v[r] = g.beta != 0.0 ? g.alpha * (double)acc[x][y][r] + g.beta * v[r] : g.alpha * (double)acc[x][y][r];
                    if (shifted) v[r] = v[r] - shift;
                }
                store_out4<TO>(C, ci, g.lo_c, v);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int er = TOUT ? rr : rr + r, ec = TOUT ? cc + r : cc;
                    if (er < g.M && ec < g.N) {
                        double v = g.alpha * (double)acc[x][y][r];
                        if (g.beta != 0.0) v = v + g.beta * load_out<TO>(C, ci + r, g.lo_c);
                        if (shifted) v = v - shift;
                        store_out<TO>(C, ci + r, g.lo_c, v);
                    }
                }
            }
        }
}

template <class TA, class TB, class TO, int BM, int BN, bool TOUT, int BK = 32>
__device__ __forceinline__ void gemm_bf16x3_tn_tile(const GemmCall &g, Bf16TnLds<BM, BN, BK> &L, int mt, int nt,
                                                    long long b)
{
    constexpr int WM = 2, WN = 2;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TM = WTM / 16, TN = WTN / 16;
    constexpr int GR = BK / 8, ROWS = kGemmThreads / GR;  // threads per staged row, rows per pass
    constexpr int RA = (BM + ROWS - 1) / ROWS, RB = (BN + ROWS - 1) / ROWS;
    static_assert(BK == 32 || BK == 64, "one or two MFMA K-slices per stage");
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int m0 = mt * BM, n0 = nt * BN;
    const int row = t / GR, grp = t % GR;
    const void *A = offset_elems<TA>(g.A, b * g.sa);
    const void *B = offset_elems<TB>(g.B, b * g.sb);
    f32x4 acc[TM][TN];
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) acc[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
    Raw8<TA> ra[RA];
    Raw8<TB> rb[RB];
    // Global loads of K-slice s+1 are issued before the MFMAs of slice s.
    auto load = [&](int k0) {
        const int kb = k0 + 8 * grp;
#pragma unroll
        for (int e = 0; e < RA; ++e) {
            const int m = row + ROWS * e;
            const bool mok = m < BM && m0 + m < g.M;
            load_raw8<TA>(A, (long long)(m0 + m) * g.lda + kb, g.lo_a, mok, g.K - kb, ra[e]);
        }
#pragma unroll
        for (int e = 0; e < RB; ++e) {
            const int n = row + ROWS * e;
            const bool nok = n < BN && n0 + n < g.N;
            load_raw8<TB>(B, kb + (long long)(n0 + n) * g.ldb, g.lo_b, nok, g.K - kb, rb[e]);
        }
    };
    load(0);
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < RA; ++e) {
            const int m = row + ROWS * e;
            if (m < BM) store_raw8<TA>(ra[e], &L.Ah[m][TnRow<BK>::at(m, grp)], &L.Al[m][TnRow<BK>::at(m, grp)]);
        }
#pragma unroll
        for (int e = 0; e < RB; ++e) {
            const int n = row + ROWS * e;
            if (n < BN) store_raw8<TB>(rb[e], &L.Bh[n][TnRow<BK>::at(n, grp)], &L.Bl[n][TnRow<BK>::at(n, grp)]);
        }
        __syncthreads();
        if (k0 + BK < g.K) load(k0 + BK);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
        bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
        for (int x = 0; x < TM; ++x) {
            const int r = wm * WTM + x * 16 + lane % 16, at = TnRow<BK>::at(r, ks / 8 + lane / 16);
            ah[x] = *reinterpret_cast<const bf16x8 *>(&L.Ah[r][at]);
            al[x] = *reinterpret_cast<const bf16x8 *>(&L.Al[r][at]);
        }
#pragma unroll
        for (int y = 0; y < TN; ++y) {
            const int c = wn * WTN + y * 16 + lane % 16, at = TnRow<BK>::at(c, ks / 8 + lane / 16);
            bh[y] = *reinterpret_cast<const bf16x8 *>(&L.Bh[c][at]);
            bl[y] = *reinterpret_cast<const bf16x8 *>(&L.Bl[c][at]);
        }
        // Each lane ends up with four output elements that are adjacent in memory. Column-major C: the
        // accumulator tile is C (lane l holds rows 4(l/16) .. +3 of column l%16). TOUT (C stored
        // transposed): the operands are swapped, the tile is C^T and lane l holds columns 4(l/16) .. +3 of
        // row l%16.
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
            for (int y = 0; y < TN; ++y) {
                if constexpr (!TOUT) {
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                } else {
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[y], al[x], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[y], ah[x], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[y], ah[x], acc[x][y], 0, 0, 0);
                }
            }
        }
    }
    tn_store_tile<TO, TM, TN, WTM, WTN, TOUT>(g, acc, m0, n0, wm, wn, lane, b);
}

template <class TA, class TB, class TO, int BM, int BN, bool TOUT, int BK = 32>
__global__ __launch_bounds__(kGemmThreads) __attribute__((amdgpu_waves_per_eu(gemm_small_waves(BM)))) void
gemm_bf16x3_tn_kernel(const GemmCall g)
{
    __shared__ Bf16TnLds<BM, BN, BK> L;
    int mt, nt;
    long long b;
    gemm_block_tile(mt, nt, b);
    gemm_bf16x3_tn_tile<TA, TB, TO, BM, BN, TOUT, BK>(g, L, mt, nt, b);
}

}  // namespace ftc
