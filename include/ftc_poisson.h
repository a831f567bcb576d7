/* Pressure Poisson solve for MI355X: the device provider that replaces the reference's poisson-solver
 * submodule, and the ELMM-side adapter that runs it inside a fort_scope external transaction.
 *
 * The reference calls the provider only through benchmarks/elmm/sources/gpu_pressure.cpp (:74, :107-108,
 * :131-132); the submodule itself (poisson-solver @ 4a1f1bb, VkFFT @ 066a17c) is not in the reference, so
 * the provider's structs below are restated from those call sites: the fields the adapter and
 * benchmarks/elmm/templates/pressure-solver.f90:16-24 read, with the same meaning.
 *
 * Problem: the FD2 7-point Laplacian of phi equals rhs on sizes[0] x sizes[1] x sizes[2] cells of the box
 * lengths[], with boundary condition pairs boundary_conditions[2d], [2d+1] per axis: 0 periodic or
 * 4 Neumann on the staggered (cell-face) boundary. Supported: x and y periodic or Neumann, z Neumann.
 * The singular (all-Neumann/periodic) problem is solved for the zero-mean phi; rhs must have zero mean.
 * Layouts: phi is dense rank 3 with 2 ghost layers (extent n+4), rhs with 1 (extent n+2), column-major;
 * only the interiors are read (rhs) and written (phi).
 *
 * Method (both precisions): the x and y directions are diagonalised by their orthonormal real
 * eigenvector bases (dense transforms run as MFMA contractions), and each (kx, ky) mode leaves an FP64
 * tridiagonal system in z solved by the Thomas algorithm.
 *   flags 0                       FP64 transforms (FP64 MFMA): one direct solve.
 *   POISSON_SOLVER_MIXED          transforms as split-bf16 MFMA products with FP32 accumulation on FP32
 *                                 workspaces, the z solves in FP64, then FP64 residuals and FP64
 *                                 iterative refinement until the relative residual reaches the tolerance.
 *   POISSON_SOLVER_REAL32         fields are REAL32 (reference flag; not supported by this provider).
 */
#ifndef FTC_POISSON_H
#define FTC_POISSON_H
#include <stddef.h>
#include <stdint.h>

#include "fort_scope.h"
#ifdef __cplusplus
extern "C" {
#endif

enum poisson_solver_status {
    POISSON_SOLVER_OK = 0, POISSON_SOLVER_ARGUMENT = 1, POISSON_SOLVER_UNSUPPORTED = 2,
    POISSON_SOLVER_RESOURCE = 3, POISSON_SOLVER_EXECUTION = 4
};
enum poisson_solver_flags { POISSON_SOLVER_REAL32 = 1, POISSON_SOLVER_MIXED = 2 };
enum poisson_solver_workspace_flags {
    POISSON_SOLVER_DEVICE_BUFFERS_EXACT = 1, POISSON_SOLVER_DRIVER_OVERHEAD_EXCLUDED = 2
};

/* gpu_pressure.cpp:51-69 reads version, flags, dimensions, approximation, sizes, lengths,
 * boundary_conditions and device_budget_bytes. */
typedef struct poisson_solver_config_v1 {
    uint32_t version, flags;
    int32_t dimensions, approximation;
    int32_t sizes[3];
    double lengths[3];
    int32_t boundary_conditions[6];
    uint64_t device_budget_bytes;
} poisson_solver_config_v1;
/* gpu_pressure.cpp:131-134 reads version, flags, persistent_device_bytes, peak_device_bytes and
 * execution_extra_bytes. */
typedef struct poisson_solver_workspace_v1 {
    uint32_t version, flags;
    uint64_t persistent_device_bytes, peak_device_bytes, execution_extra_bytes;
} poisson_solver_workspace_v1;
/* What the last execute did (no reference counterpart). */
typedef struct poisson_solver_report_v1 {
    uint32_t version, flags;
    int32_t iterations;           /* direct solves applied (1 + refinement corrections) */
    int32_t converged;            /* 1 if the relative residual reached the tolerance */
    double relative_residual;     /* ||rhs - A phi||_2 / ||rhs||_2, fixed-order FP64 sums; -1 if not measured */
    double rhs_norm;
} poisson_solver_report_v1;
typedef struct poisson_solver_handle_v1 poisson_solver_handle_v1;

int poisson_solver_create_v1(const poisson_solver_config_v1 *config, poisson_solver_handle_v1 **solver); /* gpu_pressure.cpp:131 */
int poisson_solver_workspace_query_v1(poisson_solver_handle_v1 *solver, poisson_solver_workspace_v1 *bytes); /* :132 */
/* Enqueues the solve on `stream` (a hipStream_t). *started is set to 1 before the first enqueue and 0 on a
 * rejection before any work. phi_ghosts/rhs_ghosts give the ghost widths per axis ({2,2,2} and {1,1,1}). */
int poisson_solver_execute_device_v1(poisson_solver_handle_v1 *solver, void *phi, size_t phi_bytes, const void *rhs,
                                     size_t rhs_bytes, const int32_t phi_ghosts[3], const int32_t rhs_ghosts[3],
                                     void *stream, int *started); /* gpu_pressure.cpp:107-108 */
int poisson_solver_destroy_v1(poisson_solver_handle_v1 **solver); /* gpu_pressure.cpp:74 */
const char *poisson_solver_last_error_v1(void);                   /* gpu_pressure.cpp:44 */
/* Refinement controls of the mixed solve (no reference counterpart): stop at relative residual <= tol
 * or after max_solves direct solves; measure=0 skips every residual (a fixed number of solves, no host
 * synchronisation, so the execute can be captured in a graph). Defaults: 1e-12, 6, 1. */
int poisson_solver_set_refinement_v1(poisson_solver_handle_v1 *solver, double tol, int32_t max_solves, int32_t measure);
int poisson_solver_last_report_v1(poisson_solver_handle_v1 *solver, poisson_solver_report_v1 *report);

/* ELMM adapter: benchmarks/elmm/sources/gpu_pressure.h:13-30. */
typedef struct elmm_gpu_pressure_config_v1 {
    uint32_t version, reserved;
    poisson_solver_config_v1 solver;
    int64_t phi_lower_bounds[3], rhs_lower_bounds[3];
} elmm_gpu_pressure_config_v1;
typedef struct elmm_gpu_pressure_handle_v1 elmm_gpu_pressure_handle_v1;
int elmm_gpu_pressure_create_v1(fort_scope_domain_t domain, const elmm_gpu_pressure_config_v1 *config,
                                elmm_gpu_pressure_handle_v1 **handle);        /* gpu_pressure.h:22-23 */
int elmm_gpu_pressure_solve_v1(fort_scope_t context, elmm_gpu_pressure_handle_v1 *handle, fort_buffer_t phi,
                               fort_buffer_t rhs, fort_scope_external_report_v1 *report); /* gpu_pressure.h:26-27 */
int elmm_gpu_pressure_destroy_v1(elmm_gpu_pressure_handle_v1 **handle);    /* gpu_pressure.h:29 */
const char *elmm_gpu_pressure_error_v1(void);                               /* gpu_pressure.h:30 */
/* The provider behind an adapter handle, for refinement controls and reports (no reference counterpart). */
poisson_solver_handle_v1 *elmm_gpu_pressure_provider_v1(elmm_gpu_pressure_handle_v1 *handle);

/* The nests around the solve of one domain (ELMM's prepoisson and postpoisson, benchmarks/elmm/elmm.toml:212-225,
 * whose sources are absent; the arithmetic is oracle/step.c's). U, V, W and RHS have extents (n+2), Phi (n+4),
 * column-major, as benchmarks/elmm/sources/gpu_pressure.h:8-12 lays them out.
 *   divergence:  RHS = ((dU/dx + dV/dy) + dW/dz) / dt on the interior cells (backward differences); RHS's
 *                ghosts are not written.
 *   projection:  U -= dt/dx (phi(i+1) - phi(i)), V likewise, W on the interior faces k < nz, with phi periodic
 *                in x and y (its own cells, not its ghosts); then the halo images of U, V, W as the LES step
 *                writes them (periodic x, y; U, V mirrored with a sign change across the z walls; W's walls
 *                not written). */
int cpp_fort_shared_pressure_divergence(fort_scope_t ctx, int mode, fort_buffer_t rhs, fort_buffer_t u,
                                        fort_buffer_t v, fort_buffer_t w, const double *dx, const double *dy,
                                        const double *dz, const double *dt, const int *nx, const int *ny,
                                        const int *nz);
int cpp_fort_shared_pressure_projection(fort_scope_t ctx, int mode, fort_buffer_t u, fort_buffer_t v,
                                        fort_buffer_t w, fort_buffer_t phi, const double *dx, const double *dy,
                                        const double *dz, const double *dt, const int *nx, const int *ny,
                                        const int *nz);
/* The projection with the solve's x and y boundary codes (poisson_solver_config_v1.boundary_conditions[0]
 * and [2]: 0 periodic, 4 Neumann). The projection implements periodic x and y only: a Neumann wall in x or
 * y is refused with BOUNDARY before any work, so a walled solve is never paired with periodic face
 * differences. No counterpart in the reference beyond the projection it checks. */
int cpp_fort_shared_pressure_projection_bc(fort_scope_t ctx, int mode, fort_buffer_t u, fort_buffer_t v,
                                           fort_buffer_t w, fort_buffer_t phi, const double *dx, const double *dy,
                                           const double *dz, const double *dt, const int *nx, const int *ny,
                                           const int *nz, const int *bc_x, const int *bc_y);
/* ELMM's fixpressuretoreference: phi -= *phi_ref on the interior cells of Phi (extents n+4), the cells the
 * solve defines; the ghost layers keep their values (the projection reads none of them, and a halo fill of
 * phi refills them from the shifted interior). The caller forms *phi_ref as the mean of the reference
 * plane, which templates/serial-pressure-reference.f90 licenses as a free-order sum on the device:
 * cpp_fort_shared_fixed_order_reduction over the plane, divided by nx*ny. */
int cpp_fort_shared_pressure_fix_reference(fort_scope_t ctx, int mode, fort_buffer_t phi, const double *phi_ref,
                                           const int *nx, const int *ny, const int *nz);
/* The same shift with the reference-plane sum read on the device: phi -= ref_sum(1) / (nx*ny), where
 * ref_sum is a registered REAL64 array that cpp_fort_shared_fixed_order_reduction_device filled. With
 * both queued, no value crosses to the host inside the step, so the step can replay as one graph. */
int cpp_fort_shared_pressure_fix_reference_sum(fort_scope_t ctx, int mode, fort_buffer_t phi, fort_buffer_t ref_sum,
                                               const int *nx, const int *ny, const int *nz);

#ifdef __cplusplus
}
#endif
#endif
