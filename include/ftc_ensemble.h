/* Batched ensembles of uneven LES subdomains.
 *
 * The reference merges independent loop nests into one launch when their sections are disjoint
 * (compiler/emission/cuda/kernels.py:125-160, fort_scope_graph_enqueue_members_v1 at
 * compiler/runtime/scoped_runtime.h:401-413). This object applies the same idea to whole time steps of many
 * subdomains: every stage of the step (LES tendency and update with halos, divergence, the pressure solve's
 * transforms, z solves and residuals, projection) runs for all members in one persistent launch over a
 * work list, and the whole step is one captured hipGraph. No reference counterpart beyond that analogue.
 *
 * One step of a member (fields U, V, W, S of (nx+2, ny+2, nz+2) doubles, one ghost layer):
 *   F* = F + dt * (advection + Smagorinsky diffusion)      with halos   (cpp_fort_shared_les_step)
 *   rhs = div(U*) / dt;  lap(phi) = rhs (zero-mean phi)                  (the pressure provider)
 *   U = U* - dt * grad(phi)  on the interior faces, with halos           (projection)
 * Modes:
 *   FTC_ENSEMBLE_FP64_PER_MEMBER  plain FP64, one launch per stage per member, no graph (the baseline);
 *   FTC_ENSEMBLE_FP64_BATCHED     the same FP64 arithmetic in persistent batched launches, one graph;
 *   FTC_ENSEMBLE_MIXED_BATCHED    batched; the pressure solve's transforms as split-bf16 MFMA products on
 *                                 FP32 workspaces, FP64 z solves, FP64 residual and refinement (a fixed
 *                                 number of solves, default 3), FP64 mean removal; one graph.
 * All device state lives in HBM owned by the object; upload/download copy whole member fields.
 */
#ifndef FTC_ENSEMBLE_H
#define FTC_ENSEMBLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum ftc_ensemble_mode {
    FTC_ENSEMBLE_FP64_PER_MEMBER = 0, FTC_ENSEMBLE_FP64_BATCHED = 1, FTC_ENSEMBLE_MIXED_BATCHED = 2
};
typedef struct ftc_ensemble ftc_ensemble;
typedef struct ftc_ensemble_info_v1 {
    uint32_t version, mode;
    int32_t members, solves;
    int32_t les_kchunk, gemm_tile, blocks_per_cu;
    int32_t les_policy;            /* LES tile widths: 0 all 64, 1 all 32, 2 per member by idle lanes */
    int64_t launches_per_step;     /* kernel launches one step issues (graph nodes when batched) */
    int64_t cells;                 /* interior cells of all members */
    double algorithmic_bytes;      /* FP64 step floor: sum over members of 8*(4P + 4I) + pressure */
    uint64_t device_bytes;
} ftc_ensemble_info_v1;

/* sizes: count triples (nx, ny, nz); every nx, ny >= 2, nz >= 2. dx, dy, dz: cell size; dt time step;
 * cs the Smagorinsky constant. */
int ftc_ensemble_create_v1(int count, const int32_t *sizes, double dx, double dy, double dz, double dt, double cs,
                           int mode, ftc_ensemble **out);
int ftc_ensemble_upload_v1(ftc_ensemble *e, int member, const double *u, const double *v, const double *w,
                           const double *s);
int ftc_ensemble_download_v1(ftc_ensemble *e, int member, double *u, double *v, double *w, double *s);
/* The member's last pressure (dense nx*ny*nz, zero mean). */
int ftc_ensemble_download_pressure_v1(ftc_ensemble *e, int member, double *phi);
/* Number of direct solves per pressure solve in the mixed mode (>= 1). */
int ftc_ensemble_set_solves_v1(ftc_ensemble *e, int solves);
/* on = 1 (the default): a step is the projected step above. on = 0: a step is the LES stage alone
 * (F* = F + dt * tendency, with halos), so that stage can be checked and timed by itself. */
int ftc_ensemble_set_projection_v1(ftc_ensemble *e, int on);
/* Sets the LES tile-width policy (see ftc_ensemble_info_v1.les_policy) without tuning. */
int ftc_ensemble_set_les_policy_v1(ftc_ensemble *e, int policy);
/* Picks the LES tile-width policy and k chunk (together, on the LES stage), the contraction tile and the
 * persistent grid by timing candidates on this ensemble (fixed candidate order, minimum of the median of
 * 3, ties to the first). The LES tiles are 64 x 4 or 32 x 8 columns; the policies give every member 64,
 * every member 32, or each member the width that leaves fewer idle lanes in its last tile column. With a
 * cache path, a stored choice for the same sizes and mode is reused and new choices are stored, with the
 * width each member takes. */
int ftc_ensemble_autotune_v1(ftc_ensemble *e, const char *cache_path);
/* Enqueues `steps` steps on the ensemble's stream (graph replays when batched). */
int ftc_ensemble_step_v1(ftc_ensemble *e, int steps);
int ftc_ensemble_synchronize_v1(ftc_ensemble *e);
int ftc_ensemble_stream_v1(ftc_ensemble *e, void **stream);
int ftc_ensemble_get_info_v1(ftc_ensemble *e, ftc_ensemble_info_v1 *info);
int ftc_ensemble_destroy_v1(ftc_ensemble **e);
const char *ftc_ensemble_error_v1(void);

#ifdef __cplusplus
}
#endif
#endif
