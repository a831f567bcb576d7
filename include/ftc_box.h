/* Entries of the box application's device procedures (compiler/tests/apps/box), in the shape of the
 * reference's generated entries (compiler/emission/cuda/scoped.py:237-244): context, mode, array handles,
 * then scalars by reference. The arrays are a(-1:n+2, -1:n+2, -1:n+2) REAL64 with two ghost layers; n is
 * taken from the layout, and a box smaller than 2 cells in a direction returns FORT_SCOPE_DECLINED, so the
 * procedure's own statements run. */
#ifndef FTC_BOX_H
#define FTC_BOX_H
#include "fort_scope.h"
#ifdef __cplusplus
extern "C" {
#endif

/* boundaries::periodic(a), compiler/tests/apps/box/boundaries.f90:6-41: two ghost layers copied from their
 * periodic images, x faces over the interior rows, then y faces over the whole x range, then z faces over
 * the whole x-y range (three queued passes in that order). */
int cpp_fort_shared_boundaries_periodic(fort_scope_t ctx, int mode, fort_buffer_t a);

/* stepper::diffuse(a, b), stepper.f90:8-30: b = a + dt*nu*laplacian(a) on the interior, with the module's
 * dx, dt and nu passed by the spliced call. */
int cpp_fort_shared_stepper_diffuse(fort_scope_t ctx, int mode, fort_buffer_t a, fort_buffer_t b, const double *dx,
                                    const double *dt, const double *nu);

/* stepper::store(a, b), stepper.f90:32-36: a(1:n, 1:n, 1:n) = b(1:n, 1:n, 1:n). */
int cpp_fort_shared_stepper_store(fort_scope_t ctx, int mode, fort_buffer_t a, fort_buffer_t b);

#ifdef __cplusplus
}
#endif
#endif
