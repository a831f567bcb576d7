/* Entries of the loop nests of the relax application's main program (compiler/tests/apps/relax/relax.f90),
 * in the shape of the reference's generated entries (compiler/emission/cuda/scoped.py:237-244): context,
 * mode, array handles, then scalars by reference. The reference's compiler cuts these nests out of the
 * program into generated entries named by a hash of the nest; here they are hand-written, and the
 * application builder splices a nest onto one when the hash of its normalised text is the one listed for
 * it. The arrays are u(0:n+1, 0:n+1) and v(0:n+1, 0:n+1) REAL64; an array that does not cover that range,
 * or n < 1, returns FORT_SCOPE_DECLINED, so the nest's own statements run. */
#ifndef FTC_RELAX_H
#define FTC_RELAX_H
#include "fort_scope.h"
#ifdef __cplusplus
extern "C" {
#endif

/* relax.f90:29-35: v(i,j) = 0.25*(u(i-1,j) + u(i+1,j) + u(i,j-1) + u(i,j+1)) for 1 <= i, j <= n. */
int cpp_fort_shared_relax_sweep(fort_scope_t ctx, int mode, fort_buffer_t v, fort_buffer_t u, const int *n);

/* relax.f90:36-43: change = max(change, |v(i,j) - u(i,j)|) over 1 <= i, j <= n, the value back on the host
 * when the call returns (the queued work before it runs first). */
int cpp_fort_shared_relax_change(fort_scope_t ctx, int mode, fort_buffer_t v, fort_buffer_t u, const int *n,
                                 double *change);

/* relax.f90:44: u(1:n, 1:n) = v(1:n, 1:n). */
int cpp_fort_shared_relax_copy(fort_scope_t ctx, int mode, fort_buffer_t u, fort_buffer_t v, const int *n);

#ifdef __cplusplus
}
#endif
#endif
