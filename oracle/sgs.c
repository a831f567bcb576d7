/* ORACLE — test infrastructure only (see oracle.h).
 *
 * Smagorinsky subgrid viscosity, eddy-viscosity diffusion and the explicit LES step built from them.
 *
 * The reference selects ELMM's sigma-model and momentumdiffusion_nobranch_2ord
 * (benchmarks/elmm/elmm.toml:174-177) but their sources are not part of the reference, so the nests below
 * are this repository's own statement of the configuration "advection plus Smagorinsky diffusion (Cs=0.17)":
 *
 *   nu_t = (Cs*Delta)^2 * sqrt(2 S_ij S_ij) at cell centres, Delta = (dx*dy*dz)^(1/3),
 *   S_ij from centred differences of the staggered velocities;
 *   momentum diffusion d/dx_j (nu_t d u_i / dx_j), with nu_t averaged to faces and edges;
 *   scalar diffusion d/dx_j (nu_t d s / dx_j) (turbulent Prandtl number 1).
 *
 * nu_t outside the interior is taken from its image: periodic in x and y, zero-gradient in z. The 7-point
 * constant-coefficient limit is the box diffusion of compiler/tests/apps/box/stepper.f90:8-30.
 *
 * Pinned by (tests/test_oracle.py): nu_t = 0 for a uniform wind; nu_t = (Cs*Delta)^2*gamma exactly for the
 * linear shear u = gamma*z away from the walls; nu_t >= 0; with nu_t forced constant the momentum diffusion
 * equals the box 7-point Laplacian (compiler/tests/apps/box/stepper.f90:8-15) to rounding; the diffusion
 * of a uniform scalar is zero; the scalar diffusion conserves the sum. Parity with a native build of ELMM is
 * unpinned: its sources are not in the reference.
 */
#include <math.h>
#include "oracle.h"

typedef struct {
    double rdx, rdy, rdz;    /* 1/d */
    double qdx, qdy, qdz;    /* 0.25/d */
    double rdx2, rdy2, rdz2; /* 1/(d*d) */
    double cs2d2;            /* (Cs*Delta)^2 */
} orc_sgs_coef;

/* The coefficients are formed once, on the host, and passed to the device unchanged. */
void orc_sgs_coefficients(double dx, double dy, double dz, double cs, double *out10)
{
    double delta = cbrt(dx * dy * dz);
    out10[0] = 1.0 / dx;
    out10[1] = 1.0 / dy;
    out10[2] = 1.0 / dz;
    out10[3] = 0.25 / dx;
    out10[4] = 0.25 / dy;
    out10[5] = 0.25 / dz;
    out10[6] = 1.0 / (dx * dx);
    out10[7] = 1.0 / (dy * dy);
    out10[8] = 1.0 / (dz * dz);
    out10[9] = (cs * delta) * (cs * delta);
}

static orc_sgs_coef unpack(const double *c)
{
    orc_sgs_coef r = {c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9]};
    return r;
}

#define u(a, b, c) U[ORC_IX(a, b, c)]
#define v(a, b, c) V[ORC_IX(a, b, c)]
#define w(a, b, c) W[ORC_IX(a, b, c)]

/* nu_t at an interior cell (i, j, k) in 2..n+1. */
static double nu_cell(const double *U, const double *V, const double *W, const orc_sgs_coef *q, int i, int j,
                      int k, int e1, int e2)
{
    double s11 = (u(i, j, k) - u(i - 1, j, k)) * q->rdx;
    double s22 = (v(i, j, k) - v(i, j - 1, k)) * q->rdy;
    double s33 = (w(i, j, k) - w(i, j, k - 1)) * q->rdz;
    double dudy = ((u(i, j + 1, k) + u(i - 1, j + 1, k)) - (u(i, j - 1, k) + u(i - 1, j - 1, k))) * q->qdy;
    double dvdx = ((v(i + 1, j, k) + v(i + 1, j - 1, k)) - (v(i - 1, j, k) + v(i - 1, j - 1, k))) * q->qdx;
    double dudz = ((u(i, j, k + 1) + u(i - 1, j, k + 1)) - (u(i, j, k - 1) + u(i - 1, j, k - 1))) * q->qdz;
    double dwdx = ((w(i + 1, j, k) + w(i + 1, j, k - 1)) - (w(i - 1, j, k) + w(i - 1, j, k - 1))) * q->qdx;
    double dvdz = ((v(i, j, k + 1) + v(i, j - 1, k + 1)) - (v(i, j, k - 1) + v(i, j - 1, k - 1))) * q->qdz;
    double dwdy = ((w(i, j + 1, k) + w(i, j + 1, k - 1)) - (w(i, j - 1, k) + w(i, j - 1, k - 1))) * q->qdy;
    double s12 = 0.5 * (dudy + dvdx);
    double s13 = 0.5 * (dudz + dwdx);
    double s23 = 0.5 * (dvdz + dwdy);
    double ss = 2.0 * ((s11 * s11 + s22 * s22) + s33 * s33) + 4.0 * ((s12 * s12 + s13 * s13) + s23 * s23);
    return q->cs2d2 * sqrt(ss);
}

/* nu_t over the whole padded box (1..n+2 in each axis): interior cells computed, ghosts from their image
 * (periodic in x and y, zero-gradient in z). */
void orc_sgs_nu(double *NU, const double *U, const double *V, const double *W, const double *coef, int nx,
                int ny, int nz, int e1, int e2, int e3)
{
    (void)e3;
    orc_sgs_coef q = unpack(coef);
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) NU[ORC_IX(i, j, k)] = nu_cell(U, V, W, &q, i, j, k, e1, e2);
#pragma omp parallel for
    for (int k = 1; k <= nz + 2; ++k)
        for (int j = 1; j <= ny + 2; ++j)
            for (int i = 1; i <= nx + 2; ++i) {
                int ii = i == 1 ? nx + 1 : (i == nx + 2 ? 2 : i);
                int jj = j == 1 ? ny + 1 : (j == ny + 2 ? 2 : j);
                int kk = k < 2 ? 2 : (k > nz + 1 ? nz + 1 : k);
                if (ii != i || jj != j || kk != k) NU[ORC_IX(i, j, k)] = NU[ORC_IX(ii, jj, kk)];
            }
}

#define NU(a, b, c) N[ORC_IX(a, b, c)]
#define EXY(a, b, c) (0.25 * (((NU(a, b, c) + NU(a + 1, b, c)) + NU(a, b + 1, c)) + NU(a + 1, b + 1, c)))
#define EXZ(a, b, c) (0.25 * (((NU(a, b, c) + NU(a + 1, b, c)) + NU(a, b, c + 1)) + NU(a + 1, b, c + 1)))
#define EYZ(a, b, c) (0.25 * (((NU(a, b, c) + NU(a, b + 1, c)) + NU(a, b, c + 1)) + NU(a, b + 1, c + 1)))

/* Diffusion terms at one point; N holds nu_t on the padded box. */
double orc_diff_u(const double *N, const double *U, const orc_sgs_coef *q, int i, int j, int k, int e1, int e2)
{
    double fxp = NU(i + 1, j, k) * (u(i + 1, j, k) - u(i, j, k));
    double fxm = NU(i, j, k) * (u(i, j, k) - u(i - 1, j, k));
    double fyp = EXY(i, j, k) * (u(i, j + 1, k) - u(i, j, k));
    double fym = EXY(i, j - 1, k) * (u(i, j, k) - u(i, j - 1, k));
    double fzp = EXZ(i, j, k) * (u(i, j, k + 1) - u(i, j, k));
    double fzm = EXZ(i, j, k - 1) * (u(i, j, k) - u(i, j, k - 1));
    return ((fxp - fxm) * q->rdx2 + (fyp - fym) * q->rdy2) + (fzp - fzm) * q->rdz2;
}

double orc_diff_v(const double *N, const double *V, const orc_sgs_coef *q, int i, int j, int k, int e1, int e2)
{
    double fxp = EXY(i, j, k) * (v(i + 1, j, k) - v(i, j, k));
    double fxm = EXY(i - 1, j, k) * (v(i, j, k) - v(i - 1, j, k));
    double fyp = NU(i, j + 1, k) * (v(i, j + 1, k) - v(i, j, k));
    double fym = NU(i, j, k) * (v(i, j, k) - v(i, j - 1, k));
    double fzp = EYZ(i, j, k) * (v(i, j, k + 1) - v(i, j, k));
    double fzm = EYZ(i, j, k - 1) * (v(i, j, k) - v(i, j, k - 1));
    return ((fxp - fxm) * q->rdx2 + (fyp - fym) * q->rdy2) + (fzp - fzm) * q->rdz2;
}

double orc_diff_w(const double *N, const double *W, const orc_sgs_coef *q, int i, int j, int k, int e1, int e2)
{
    double fxp = EXZ(i, j, k) * (w(i + 1, j, k) - w(i, j, k));
    double fxm = EXZ(i - 1, j, k) * (w(i, j, k) - w(i - 1, j, k));
    double fyp = EYZ(i, j, k) * (w(i, j + 1, k) - w(i, j, k));
    double fym = EYZ(i, j - 1, k) * (w(i, j, k) - w(i, j - 1, k));
    double fzp = NU(i, j, k + 1) * (w(i, j, k + 1) - w(i, j, k));
    double fzm = NU(i, j, k) * (w(i, j, k) - w(i, j, k - 1));
    return ((fxp - fxm) * q->rdx2 + (fyp - fym) * q->rdy2) + (fzp - fzm) * q->rdz2;
}

double orc_diff_s(const double *N, const double *S, const orc_sgs_coef *q, int i, int j, int k, int e1, int e2)
{
#define s(a, b, c) S[ORC_IX(a, b, c)]
    double fxp = (0.5 * (NU(i, j, k) + NU(i + 1, j, k))) * (s(i + 1, j, k) - s(i, j, k));
    double fxm = (0.5 * (NU(i - 1, j, k) + NU(i, j, k))) * (s(i, j, k) - s(i - 1, j, k));
    double fyp = (0.5 * (NU(i, j, k) + NU(i, j + 1, k))) * (s(i, j + 1, k) - s(i, j, k));
    double fym = (0.5 * (NU(i, j - 1, k) + NU(i, j, k))) * (s(i, j, k) - s(i, j - 1, k));
    double fzp = (0.5 * (NU(i, j, k) + NU(i, j, k + 1))) * (s(i, j, k + 1) - s(i, j, k));
    double fzm = (0.5 * (NU(i, j, k - 1) + NU(i, j, k))) * (s(i, j, k) - s(i, j, k - 1));
#undef s
    return ((fxp - fxm) * q->rdx2 + (fyp - fym) * q->rdy2) + (fzp - fzm) * q->rdz2;
}

/* Adds the diffusion of U, V, W and S to the tendencies T* (T += D), with the same loop domains as the
 * advection nests: U, V and S over all interior cells, W over the interior faces k = 2..nz. */
void orc_sgs_diffusion(double *TU, double *TV, double *TW, double *TS, const double *N, const double *U,
                       const double *V, const double *W, const double *S, const double *coef, int nx, int ny,
                       int nz, int e1, int e2, int e3)
{
    (void)e3;
    orc_sgs_coef q = unpack(coef);
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) {
                size_t p = ORC_IX(i, j, k);
                TU[p] = TU[p] + orc_diff_u(N, U, &q, i, j, k, e1, e2);
                TV[p] = TV[p] + orc_diff_v(N, V, &q, i, j, k, e1, e2);
                if (k <= nz) TW[p] = TW[p] + orc_diff_w(N, W, &q, i, j, k, e1, e2);
                TS[p] = TS[p] + orc_diff_s(N, S, &q, i, j, k, e1, e2);
            }
}
#undef NU
#undef EXY
#undef EXZ
#undef EYZ
#undef u
#undef v
#undef w
