/* Checks of the intrinsic semantics generated kernels rely on (fortran-to-cuda_amd/csrc/numeric.hpp restates
 * compiler/runtime/numeric.hpp:11-112 for host and gfx950 code).
 */
#ifndef FTC_NUMERIC_H
#define FTC_NUMERIC_H
#include "fort_scope.h"
#ifdef __cplusplus
extern "C" {
#endif

/* For n host samples a[i], b[i] of width 4 (float) or 8 (double) bytes: reference[i] = a/b and
 * candidate[i] = divide(a, b) evaluated on the device, then reference[n+i] = (i-41)/b and candidate[n+i] =
 * divide(i-41, b) (an integer numerator). reference and candidate hold 2n values. The division check of
 * compiler/tests/test_exact_division.py:33-96 compares the two bit for bit. */
int ftc_divide_probe(int width, const void *a, const void *b, void *reference, void *candidate, int n);
/* Evaluates the boundary cases of compiler/tests/test_intrinsics.py:145-189 (and the ordered min/max
 * NaN and signed-zero cases) on the device; *failures receives a bitmask of the failed checks (0: all
 * pass). */
int ftc_intrinsic_probe(int *failures);

#ifdef __cplusplus
}
#endif
#endif
