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

namespace ftc {

enum GemmPrecision : int { kF64 = 0, kBf16x3 = 1 };
enum GemmType : int { kDouble = 0, kFloat = 1 };
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
};

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
