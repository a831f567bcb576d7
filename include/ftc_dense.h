/* Dense contractions on the MI355X matrix cores.
 *
 * The reference compiles dense loop nests such as PolyBench 2mm/3mm (compiler/tests/test_polybench.py:27-41,
 * the archive itself is fetched at test time and not in the reference) into generated grid-stride kernels;
 * here the contraction D = A*B*C is one entry whose two products run as MFMA tiles.
 */
#ifndef FTC_DENSE_H
#define FTC_DENSE_H
#include <stddef.h>

#include "fort_scope.h"
#ifdef __cplusplus
extern "C" {
#endif

enum ftc_gemm_precision { FTC_GEMM_F64 = 0, FTC_GEMM_BF16X3 = 1 };
enum ftc_gemm_type { FTC_GEMM_DOUBLE = 0, FTC_GEMM_FLOAT = 1 };
enum ftc_gemm_flags { FTC_GEMM_A_K_CONTIGUOUS = 1, FTC_GEMM_TRANSPOSED_C = 2 };

/* C_b = alpha * A_b * B_b + beta * C_b for b < batch, column-major, A_b = a + b*stride_a (m x k, lda),
 * B_b = b + b*stride_b (k x n, ldb), C_b = c + b*stride_c (m x n, ldc); beta == 0 does not read C.
 * flags (split-bf16 only): FTC_GEMM_A_K_CONTIGUOUS stores A(m, k) at a[m*lda + k];
 * FTC_GEMM_TRANSPOSED_C writes C(m, n) at c[m*ldc + n] (requires a K-contiguous A). */
typedef struct ftc_gemm_desc {
    const void *a;
    const void *b;
    void *c;
    int type_a, type_b, type_c;
    int m, n, k, batch;
    long long lda, ldb, ldc;
    long long stride_a, stride_b, stride_c;
    double alpha, beta;
    int flags;
} ftc_gemm_desc;

/* Device-pointer contraction on `stream` (a hipStream_t). precision FTC_GEMM_F64: FP64 MFMA, double
 * operands. FTC_GEMM_BF16X3: operands split into bf16 hi + lo on their way into LDS, lo*hi + hi*lo + hi*hi
 * accumulated in FP32 (relative error about 2^-16 per product). tile: -1 picks the default tile shape,
 * 0 .. ftc_gemm_tile_count()-1 forces one (the autotuner's search space). No reference counterpart. */
int ftc_gemm_device(const ftc_gemm_desc *desc, int precision, int tile, void *stream);
int ftc_gemm_tile_count(void);

/* D = A * B * C for n x n column-major REAL64 arrays registered in the context (the dense chain of
 * PolyBench 3mm with the middle product reused, test_polybench.py:27-41). *precision is a
 * ftc_gemm_precision. D is INTENT(OUT). */
int cpp_fort_shared_chain_contraction(fort_scope_t ctx, int mode, fort_buffer_t d, fort_buffer_t a, fort_buffer_t b,
                                      fort_buffer_t c, const int *n, const int *precision);

/* One MFMA with one-hot operands (lane < 0: all ones); writes the 64 lanes x 4 accumulator elements.
 * kind 0: v_mfma_f64_16x16x4f64, 1: v_mfma_f32_16x16x32_bf16. Checks the lane layouts the kernels
 * assume. No reference counterpart. */
int ftc_mfma_layout_probe(int kind, int a_lane, int a_elem, int b_lane, int b_elem, double *out256);

#ifdef __cplusplus
}
#endif
#endif
