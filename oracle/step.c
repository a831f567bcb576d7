/* ORACLE — test infrastructure only (see oracle.h).
 *
 * The synthetic LES tendency and explicit step on the resident path:
 *
 *   TU = CDU + DU, TV = CDV + DV, TW = CDW + DW, TS = SADV + DS   (D* = Smagorinsky diffusion, optional)
 *   F_new = F + dt * TF on the interior, then the halo fill of F_new (orc_halo_les).
 *
 * The advection passes are the reference nests (benchmarks/cdu/src/cdu.f90:12-133,
 * compiler/tests/fixtures/stencils/{cvd,cdw}.f90:12-125); W's loop domain is the interior faces k = 2..nz
 * (Wnz = nz - 1), the walls being the faces k = 1 and k = nz + 1. Pinned through its parts; the composition
 * (tendency, then F + dt*T, then the halo) is this repository's own statement of an explicit step.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void orc_cdu(double *, const double *, const double *, const double *, double, double, double, int, int, int,
             int, int, int);
void orc_cdv(double *, const double *, const double *, const double *, double, double, double, int, int, int,
             int, int, int);
void orc_cdw(double *, const double *, const double *, const double *, double, double, double, int, int, int,
             int, int, int);
void orc_scalar_adv(double *, const double *, const double *, const double *, const double *, double, double,
                    double, int, int, int, int, int, int);
void orc_sgs_nu(double *, const double *, const double *, const double *, const double *, int, int, int, int,
                int, int);
void orc_sgs_diffusion(double *, double *, double *, double *, const double *, const double *, const double *,
                       const double *, const double *, const double *, int, int, int, int, int, int);
void orc_halo_les(double *, int, int, int, int, int, int, int);

/* Tendencies over the interior; with coef == NULL the advection alone (configuration "advection entry
 * alone"). NU is caller scratch of the padded size (unused without coef). */
void orc_les_tendency(double *TU, double *TV, double *TW, double *TS, double *NU, const double *U,
                      const double *V, const double *W, const double *S, double dx, double dy, double dz,
                      const double *coef, int nx, int ny, int nz)
{
    int e1 = nx + 2, e2 = ny + 2, e3 = nz + 2;
    orc_cdu(TU, U, V, W, dx, dy, dz, nx, ny, nz, e1, e2, e3);
    orc_cdv(TV, U, V, W, dx, dy, dz, nx, ny, nz, e1, e2, e3);
    orc_cdw(TW, U, V, W, dx, dy, dz, nx, ny, nz - 1, e1, e2, e3);
    orc_scalar_adv(TS, U, V, W, S, dx, dy, dz, nx, ny, nz, e1, e2, e3);
    if (coef) {
        orc_sgs_nu(NU, U, V, W, coef, nx, ny, nz, e1, e2, e3);
        orc_sgs_diffusion(TU, TV, TW, TS, NU, U, V, W, S, coef, nx, ny, nz, e1, e2, e3);
    }
}

/* One explicit step from (U, V, W, S) into (Un, Vn, Wn, Sn). The new fields' halos are filled; their wall
 * faces for W stay as the caller initialised them (zero), exactly as orc_halo_les leaves them. */
void orc_les_step(double *Un, double *Vn, double *Wn, double *Sn, const double *U, const double *V,
                  const double *W, const double *S, double dx, double dy, double dz, double dt,
                  const double *coef, int nx, int ny, int nz)
{
    int e1 = nx + 2, e2 = ny + 2, e3 = nz + 2;
    size_t P = (size_t)e1 * e2 * e3;
    double *T = (double *)malloc(5 * P * sizeof(double));
    double *TU = T, *TV = T + P, *TW = T + 2 * P, *TS = T + 3 * P, *NU = T + 4 * P;
    memset(T, 0, 5 * P * sizeof(double));
    orc_les_tendency(TU, TV, TW, TS, NU, U, V, W, S, dx, dy, dz, coef, nx, ny, nz);
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) {
                size_t p = ORC_IX(i, j, k);
                Un[p] = U[p] + dt * TU[p];
                Vn[p] = V[p] + dt * TV[p];
                if (k <= nz) Wn[p] = W[p] + dt * TW[p];
                Sn[p] = S[p] + dt * TS[p];
            }
    orc_halo_les(Un, 0, nx, ny, nz, e1, e2, e3);
    orc_halo_les(Vn, 0, nx, ny, nz, e1, e2, e3);
    orc_halo_les(Wn, 1, nx, ny, nz, e1, e2, e3);
    orc_halo_les(Sn, 2, nx, ny, nz, e1, e2, e3);
    free(T);
}
