/* Hot-path entries of the MI355X resident LES backend.
 *
 * Every entry has the shape of the reference compiler's generated entries
 * (compiler/emission/cuda/scoped.py:237-244, ENTRY_ABI_VERSION 3): the scoped context, the execution mode
 * (0 native, 1 device), the array handles in argument order, then the scalars by reference, exactly as a
 * Fortran interface `function run(...) bind(C, name='...')` passes them (handles by value, scalars
 * without VALUE). The return value is a fort_scope status; FORT_SCOPE_DECLINED (9) means nothing was done
 * and the caller runs its original statements (native mode always declines: this library holds device
 * implementations only). Work is queued on the context and submitted by fort_scope_graph_flush_v1, a host
 * access or the queue's own threshold.
 */
#ifndef FTC_LES_H
#define FTC_LES_H
#include "fort_scope.h"
#ifdef __cplusplus
extern "C" {
#endif

/* subroutine CDU(U2, U, V, W, dxmin, dymin, dzmin, Unx, Uny, Unz): benchmarks/cdu/src/cdu.f90:112-133 (the set,
 * CDUdiv, CDUadv and multiply passes fused into one kernel, evaluated per point in source order). U2's
 * interior is overwritten (INTENT(OUT): its other values become undefined). */
int cpp_fort_shared_momentumadvection_cdu(fort_scope_t ctx, int mode, fort_buffer_t u2, fort_buffer_t u,
                                          fort_buffer_t v, fort_buffer_t w, const double *dxmin,
                                          const double *dymin, const double *dzmin, const int *unx,
                                          const int *uny, const int *unz);
/* subroutine CDV(V2, U, V, W, ...): compiler/tests/fixtures/stencils/cvd.f90:104-125. */
int cpp_fort_shared_momentumadvection_cdv(fort_scope_t ctx, int mode, fort_buffer_t v2, fort_buffer_t u,
                                          fort_buffer_t v, fort_buffer_t w, const double *dxmin,
                                          const double *dymin, const double *dzmin, const int *vnx,
                                          const int *vny, const int *vnz);
/* subroutine CDW(W2, U, V, W, ...): compiler/tests/fixtures/stencils/cdw.f90:104-125. */
int cpp_fort_shared_momentumadvection_cdw(fort_scope_t ctx, int mode, fort_buffer_t w2, fort_buffer_t u,
                                          fort_buffer_t v, fort_buffer_t w, const double *dxmin,
                                          const double *dymin, const double *dzmin, const int *wnx,
                                          const int *wny, const int *wnz);

/* LES tendencies on nx*ny*nz cells with one ghost layer (arrays of extents nx+2, ny+2, nz+2): the three
 * momentum advection components and flux-form scalar advection, plus Smagorinsky diffusion when
 * *sgs != 0 (coefficient *cs). Replaces the advection and subgrid selections of ELMM's step
 * (benchmarks/elmm/elmm.toml:171-200), whose Fortran is not in the reference; W is computed on the
 * interior faces k = 2..nz. */
int cpp_fort_shared_les_tendency(fort_scope_t ctx, int mode, fort_buffer_t tu, fort_buffer_t tv, fort_buffer_t tw,
                                 fort_buffer_t ts, fort_buffer_t u, fort_buffer_t v, fort_buffer_t w,
                                 fort_buffer_t s, const double *dx, const double *dy, const double *dz,
                                 const double *cs, const int *nx, const int *ny, const int *nz, const int *sgs);

/* One explicit LES step: F_new = F + dt * tendency on the interior, then the halo of F_new (periodic in x
 * and y; no-slip walls in z: tangential velocities mirrored with a sign change, W zero on the walls, the
 * scalar zero-gradient), all in one kernel. The walls and the top ghost of W_new are not written and keep
 * the values (zero) the caller gave them. */
int cpp_fort_shared_les_step(fort_scope_t ctx, int mode, fort_buffer_t un, fort_buffer_t vn, fort_buffer_t wn,
                             fort_buffer_t sn, fort_buffer_t u, fort_buffer_t v, fort_buffer_t w, fort_buffer_t s,
                             const double *dx, const double *dy, const double *dz, const double *dt,
                             const double *cs, const int *nx, const int *ny, const int *nz, const int *sgs);

/* The 19 host-side coefficients the LES entries pass to their kernels: 0.25/d, 0.5/d, 0.125/d for
 * d = dx, dy, dz, then 1/d, 0.25/d, 1/d^2 and (Cs*cbrt(dx*dy*dz))^2. No reference counterpart: it lets a
 * checker restate the step with the very same constants. */
void ftc_les_coefficients(double dx, double dy, double dz, double cs, double *out19);

/* HIP events on a context stream (fort_scope_gpu_enter returns the stream) for host-side timing, and the
 * visible device count. No reference counterpart. */
int ftc_event_create(void **event);
int ftc_event_record(void *event, void *stream);
int ftc_event_elapsed_ms(void *start, void *stop, float *ms);
int ftc_event_destroy(void *event);
int ftc_device_count(void);
/* A plain device-to-device copy of n doubles, launched reps times: the measured HBM ceiling and the
 * calibration kernel of the HBM byte counters. No reference counterpart. */
int ftc_copy_probe(size_t n, int reps, float *ms_per_launch);
/* The HBM ceilings of this part: kind 0 copy (the probe above), 1 copy with nontemporal loads and stores,
 * 2 read only, 3 write only, over n doubles (n even). No reference counterpart. */
int ftc_bandwidth_probe(int kind, size_t n, int reps, float *ms_per_launch);
/* The FP64 matrix-core ceiling of this part: waves_per_simd (1..8) four-wave workgroups per CU, each wave
 * iters steps of eight independent 16x16x4 FP64 MFMAs (2048 flops each, no memory traffic), launched reps
 * times after a warm-up; writes the milliseconds and the flops of one launch. No reference counterpart. */
int ftc_mfma_probe(int waves_per_simd, int iters, int reps, float *ms_per_launch, double *flops_per_launch);
/* Visit counts of the tile orders the kernels use (csrc/xcd.hpp), for a check that each covers every tile
 * exactly once (as compiler/tests/test_cuda_coordinate_quotient.py counts the visits of the reference's
 * grid-stride decode): scheme 0 the launch order and 1 the XCD-contiguous order of a gx x gy x gz grid, 2
 * the persistent banded walk of gx*gy*gz tiles by `workgroups` (a positive multiple of 8), 3 the
 * contraction order of a gx x gy x gz grid. counts has gx*gy*gz + 1 entries: the visits of each tile, then
 * of any tile number past the last (0 for a correct order). No reference counterpart. */
int ftc_tile_visit_probe_v1(int scheme, unsigned gx, unsigned gy, unsigned gz, unsigned workgroups, unsigned *counts);
/* The device function of a kernel (double *x, long long n, double scale) that writes x[i] = scale * i for
 * i < n with one thread per element (256-thread blocks): a task body for exercising the graph queue and
 * merged member launches from a caller that has no kernels of its own. No reference counterpart. */
const void *ftc_iota_kernel(void);

#ifdef __cplusplus
}
#endif
#endif
