// Dense column-major contractions on the MFMA units of gfx950, shared by the chain contraction entry
// and the spectral transforms of the pressure solve.
//
//   C_b = alpha * A_b * B_b + beta * C_b,   b = 0 .. batch-1,
//   A_b = A + b*sa (M x K, leading dimension lda), B_b = B + b*sb (K x N, ldb), C_b = C + b*sc (M x N, ldc)
//
// Precisions:
//   kF64      FP64 inputs and FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 accumulation;
//   kBf16x3   inputs split into bf16 hi + lo parts on their way into LDS; each 16x16x32 bf16 MFMA step
//             accumulates lo*hi + hi*lo + hi*hi into FP32 accumulators (about 2^-16 relative error).
// Inputs and outputs are double or float (the element type of each operand is part of the call).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace ftc {

enum GemmPrecision : int { kF64 = 0, kBf16x3 = 1 };
// kSplit: a split-bf16 operand kept as two bf16 planes, hi at the pointer and lo `lo` elements after it
// (GemmCall::lo_a / lo_b / lo_c). Such operands (and outputs) skip the split on the way into LDS; only the
// K-contiguous split-bf16 kernels take them, and only calls built inside the library.
enum GemmType : int { kDouble = 0, kFloat = 1, kSplit = 2 };
// Layout flags (split-bf16 only): A stored K-contiguous, A(m, k) = A[m*lda + k]; C written transposed,
// C(m, n) = C[m*ldc + n]. With both operands K-contiguous every tile row is one vector load.
enum GemmFlags : int { kGemmAKContiguous = 1, kGemmTransposedC = 2 };

struct GemmCall {
    const void *A;
    const void *B;
    void *C;
    int typeA, typeB, typeC;
    int M, N, K, batch;
    long long lda, ldb, ldc;
    long long sa, sb, sc;
    double alpha, beta;
    int flags;
    long long lo_a, lo_b, lo_c;  // lo-plane offsets (elements) of kSplit operands
    // Split-bf16 K-contiguous calls only: when set, every stored C value is v - *shift (a device scalar).
    const double *shift = nullptr;
};

// Host-side split of values into the kSplit layout: the value rounded to float, hi = its bf16 rounding
// (nearest even), lo = the bf16 rounding of the float remainder; out holds all hi values, then all lo.
inline uint16_t bf16_bits_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

inline float bf16_bits_to_float(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

inline std::vector<uint16_t> split_planes(const std::vector<double> &v)
{
    std::vector<uint16_t> out(2 * v.size());
    for (std::size_t e = 0; e < v.size(); ++e) {
        const float f = (float)v[e];
        const uint16_t h = bf16_bits_rne(f);
        out[e] = h;
        out[v.size() + e] = bf16_bits_rne(f - bf16_bits_to_float(h));
    }
    return out;
}

// Launch geometry chosen for a call (block tile and grid); `tile` indexes the compiled tile shapes.
struct GemmPlan {
    int tile;
    dim3 grid, block;
};

// Validates the call and enqueues it on `stream`. Returns a fort_scope status.
int gemm_launch(const GemmCall &call, int precision, hipStream_t stream, int tile = -1);
// The tile shape gemm_launch picks for this call when `tile` is -1.
int gemm_default_tile(const GemmCall &call, int precision);
int gemm_tile_count(int precision);

}  // namespace ftc
