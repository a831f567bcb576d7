/* Fixed-order reductions of resident arrays.
 *
 * The reference lowers an accumulating nest into up to 65536 slots (slot s takes iterations s, s+65536, ...
 * in order), a fold into 256 values (value f takes slots f, f+256, ...) and a host fold of the 256 in index
 * order; nests of at most 65536 iterations go straight to 256 values (compiler/resident/reductions.py:1-45,
 * constants :66-72). Max and min then equal the serial result and real sums repeat exactly run to run.
 */
#ifndef FTC_REDUCE_H
#define FTC_REDUCE_H
#include "fort_scope.h"
#ifdef __cplusplus
extern "C" {
#endif

enum ftc_reduction { FTC_REDUCE_SUM = 0, FTC_REDUCE_SUM_SQUARES = 1, FTC_REDUCE_MAX = 2, FTC_REDUCE_MIN = 3,
                     FTC_REDUCE_MAX_ABS = 4 };

/* Reduces the zero-based half-open box [lo3, hi3) of a registered rank-3 REAL64 array (iterations in
 * Fortran order over the box) and folds the result into *value (in: the variable before the nest; out:
 * after). max/min fold as `value > x ? value : x` (compiler/runtime/numeric.hpp:48-58). The result is on
 * the host when the call returns. */
int cpp_fort_shared_fixed_order_reduction(fort_scope_t ctx, int mode, fort_buffer_t x, const int *op,
                                          const long long *lo3, const long long *hi3, double *value);

/* The same reduction as tasks of the context's graph queue, as the reference lowers it
 * (compiler/resident/reductions.py:18-31): the 65,536 slot and 256 fold values go to the registered REAL64
 * work array `partial` (at least 65,792 elements along its first dimension, later dimensions of extent 1),
 * and the fold of the 256 values into *initial, in index order, to result(1) on the device. Nothing is
 * synchronised and nothing returns to the host: the value is observed like any array, or read by a later
 * entry (cpp_fort_shared_pressure_fix_reference_sum). Bit-identical to
 * cpp_fort_shared_fixed_order_reduction with *value = *initial. */
int cpp_fort_shared_fixed_order_reduction_device(fort_scope_t ctx, int mode, fort_buffer_t x, fort_buffer_t partial,
                                                 fort_buffer_t result, const int *op, const long long *lo3,
                                                 const long long *hi3, const double *initial);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under stream or
 * graph capture, serving loops, harnesses that forbid allocation): the same fixed-order reduction of the
 * box x[i + j*ld1 + k*ld2], 0 <= i < n1, 0 <= j < n2, 0 <= k < n3, of a double array the caller has in
 * device memory, folded into `initial` and stored in *result (device memory), enqueued on `stream` (a
 * hipStream_t; NULL is the null stream) with no host synchronisation. The call allocates nothing: its
 * 65,536 slot and 256 fold values live in one static device array of 526,336 bytes per device, which every
 * call on that device shares and which keeps the last call's values. Calls that overlap on one device (on
 * different streams, or from different threads) race on it and must be serialised by the caller; every
 * call overwrites all it later reads, so nothing has to be cleared. Bit-identical to
 * cpp_fort_shared_fixed_order_reduction on the same box. Boxes beyond 2^62 elements are refused with
 * BOUNDARY. */
int ftc_fixed_order_reduction_static_v1(const double *x, long long n1, long long n2, long long n3, long long ld1,
                                        long long ld2, int op, double initial, double *result, void *stream);

#ifdef __cplusplus
}
#endif
#endif
