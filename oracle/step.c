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

/* The projection half of the step (ELMM's prepoisson and postpoisson, benchmarks/elmm/elmm.toml:212-217,
 * whose sources are not in the reference; written from the FD2 operators of the pressure provider,
 * benchmarks/elmm/sources/gpu_pressure.cpp:51-69):
 *   rhs(c) = (((U(i) - U(i-1))/dx + (V(j) - V(j-1))/dy) + (W(k) - W(k-1))/dz) / dt   at every cell,
 *   U(i) -= (dt/dx) (phi(c+1) - phi(c)) on x faces (periodic), likewise V; W on the interior z faces,
 * then the halo fill of U, V (tangential) and W (wall-normal). rhs and phi are dense nx*ny*nz arrays.
 * Pinned by: the divergence of a projected field with an exact pressure is zero to rounding (div grad is
 * the 7-point Laplacian of the pressure solve), tests/test_oracle.py. */
void orc_divergence(double *rhs, const double *U, const double *V, const double *W, double dx, double dy,
                    double dz, double dt, int nx, int ny, int nz)
{
    int e1 = nx + 2, e2 = ny + 2;
    double rdx = 1.0 / dx, rdy = 1.0 / dy, rdz = 1.0 / dz, rdt = 1.0 / dt;
#pragma omp parallel for
    for (int c = 1; c <= nz; ++c)
        for (int b = 1; b <= ny; ++b)
            for (int a = 1; a <= nx; ++a) {
                size_t p = ORC_IX(a + 1, b + 1, c + 1);
                size_t q = (size_t)(a - 1) + (size_t)nx * ((size_t)(b - 1) + (size_t)ny * (size_t)(c - 1));
                rhs[q] = (((U[p] - U[p - 1]) * rdx + (V[p] - V[p - e1]) * rdy) +
                          (W[p] - W[p - (size_t)e1 * e2]) * rdz) * rdt;
            }
}

void orc_projection(double *U, double *V, double *W, const double *phi, double dx, double dy, double dz, double dt,
                    int nx, int ny, int nz)
{
    int e1 = nx + 2, e2 = ny + 2, e3 = nz + 2;
    double gx = dt / dx, gy = dt / dy, gz = dt / dz;
    size_t plane = (size_t)nx * ny;
#pragma omp parallel for
    for (int c = 1; c <= nz; ++c)
        for (int b = 1; b <= ny; ++b)
            for (int a = 1; a <= nx; ++a) {
                size_t p = ORC_IX(a + 1, b + 1, c + 1);
                size_t q = (size_t)(a - 1) + (size_t)nx * (size_t)(b - 1) + plane * (size_t)(c - 1);
                size_t qx = (size_t)(a % nx) + (size_t)nx * (size_t)(b - 1) + plane * (size_t)(c - 1);
                size_t qy = (size_t)(a - 1) + (size_t)nx * (size_t)(b % ny) + plane * (size_t)(c - 1);
                U[p] = U[p] - gx * (phi[qx] - phi[q]);
                V[p] = V[p] - gy * (phi[qy] - phi[q]);
                if (c < nz) W[p] = W[p] - gz * (phi[q + plane] - phi[q]);
            }
    orc_halo_les(U, 0, nx, ny, nz, e1, e2, e3);
    orc_halo_les(V, 0, nx, ny, nz, e1, e2, e3);
    orc_halo_les(W, 1, nx, ny, nz, e1, e2, e3);
}
