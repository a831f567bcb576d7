/* ORACLE — test infrastructure only (see oracle.h).
 *
 * Staggered-grid momentum advection (CDU, CDV, CDW) and flux-form scalar advection.
 * U(i,j,k) sits on the x face i+1/2, V(i,j,k) on the y face j+1/2, W(i,j,k) on the z face k+1/2, and the
 * scalar at the cell centre.
 */
#include "oracle.h"

/* CDU: benchmarks/cdu/src/cdu.f90:12-133 (set, CDUdiv, CDUadv, multiply, called in that order by CDU).
 * Pinned by: bit equality with the reference's own cdu.f90 compiled by the oracle Makefile
 * (oracle/_ref/libref_cdu.so) on seeded inputs, and with the golden dump of the reference driver
 * compiler/tests/fixtures/stencils/cdu_main.f90 at 16^3 (tests/golden/). */
void orc_cdu(double *U2, const double *U, const double *V, const double *W, double dxmin, double dymin,
             double dzmin, int nx, int ny, int nz, int e1, int e2, int e3)
{
    (void)e3;
    /* set(U2, 0): cdu.f90:12-27 */
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) U2[ORC_IX(i, j, k)] = 0.0;
    /* CDUdiv: cdu.f90:48-78 */
    {
        const double Ax = 0.25 / dxmin, Ay = 0.25 / dymin, Az = 0.25 / dzmin;
#pragma omp parallel for
        for (int k = 2; k <= nz + 1; ++k)
            for (int j = 2; j <= ny + 1; ++j)
                for (int i = 2; i <= nx + 1; ++i) {
#define u(a, b, c) U[ORC_IX(a, b, c)]
#define v(a, b, c) V[ORC_IX(a, b, c)]
#define w(a, b, c) W[ORC_IX(a, b, c)]
                    double x = Ax * (u(i + 1, j, k) + u(i, j, k)) * (u(i + 1, j, k) + u(i, j, k)) -
                               Ax * (u(i, j, k) + u(i - 1, j, k)) * (u(i, j, k) + u(i - 1, j, k));
                    double y = Ay * (u(i, j + 1, k) + u(i, j, k)) * (v(i + 1, j, k) + v(i, j, k)) -
                               Ay * (u(i, j, k) + u(i, j - 1, k)) * (v(i + 1, j - 1, k) + v(i, j - 1, k));
                    double z = Az * (u(i, j, k + 1) + u(i, j, k)) * (w(i + 1, j, k) + w(i, j, k)) -
                               Az * (u(i, j, k) + u(i, j, k - 1)) * (w(i + 1, j, k - 1) + w(i, j, k - 1));
                    U2[ORC_IX(i, j, k)] = -((x + y) + z);
                }
    }
    /* CDUadv: cdu.f90:80-107 */
    {
        const double Ax = 0.5 / dxmin, Ay = 0.125 / dymin, Az = 0.125 / dzmin;
#pragma omp parallel for
        for (int k = 2; k <= nz + 1; ++k)
            for (int j = 2; j <= ny + 1; ++j)
                for (int i = 2; i <= nx + 1; ++i) {
                    double Vadv = v(i, j, k) + v(i + 1, j, k) + v(i, j - 1, k) + v(i + 1, j - 1, k);
                    double Wadv = w(i, j, k) + w(i + 1, j, k) + w(i, j, k - 1) + w(i + 1, j, k - 1);
                    U2[ORC_IX(i, j, k)] = U2[ORC_IX(i, j, k)] -
                                          (Ax * (u(i + 1, j, k) - u(i - 1, j, k)) * u(i, j, k) +
                                           Ay * (u(i, j + 1, k) - u(i, j - 1, k)) * Vadv +
                                           Az * (u(i, j, k + 1) - u(i, j, k - 1)) * Wadv);
                }
    }
    /* multiply(U2, 0.5): cdu.f90:29-44 */
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) U2[ORC_IX(i, j, k)] = U2[ORC_IX(i, j, k)] * 0.5;
}

/* CDV: compiler/tests/fixtures/stencils/cvd.f90:12-125 (CDVdiv :44-71, CDVadv :73-98).
 * Pinned by: bit equality with the reference cvd.f90 built into oracle/_ref and the 16^3 golden dump of
 * cvd_main.f90. */
void orc_cdv(double *V2, const double *U, const double *V, const double *W, double dxmin, double dymin,
             double dzmin, int nx, int ny, int nz, int e1, int e2, int e3)
{
    (void)e3;
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) V2[ORC_IX(i, j, k)] = 0.0;
    {
        const double Ax = 0.25 / dxmin, Ay = 0.25 / dymin, Az = 0.25 / dzmin;
#pragma omp parallel for
        for (int k = 2; k <= nz + 1; ++k)
            for (int j = 2; j <= ny + 1; ++j)
                for (int i = 2; i <= nx + 1; ++i) {
                    double y = Ay * (v(i, j + 1, k) + v(i, j, k)) * (v(i, j + 1, k) + v(i, j, k)) -
                               Ay * (v(i, j, k) + v(i, j - 1, k)) * (v(i, j, k) + v(i, j - 1, k));
                    double x = Ax * (v(i + 1, j, k) + v(i, j, k)) * (u(i, j + 1, k) + u(i, j, k)) -
                               Ax * (v(i, j, k) + v(i - 1, j, k)) * (u(i - 1, j + 1, k) + u(i - 1, j, k));
                    double z = Az * (v(i, j, k + 1) + v(i, j, k)) * (w(i, j + 1, k) + w(i, j, k)) -
                               Az * (v(i, j, k) + v(i, j, k - 1)) * (w(i, j + 1, k - 1) + w(i, j, k - 1));
                    V2[ORC_IX(i, j, k)] = -((y + x) + z);
                }
    }
    {
        const double Ax = 0.125 / dxmin, Ay = 0.5 / dymin, Az = 0.125 / dzmin;
#pragma omp parallel for
        for (int k = 2; k <= nz + 1; ++k)
            for (int j = 2; j <= ny + 1; ++j)
                for (int i = 2; i <= nx + 1; ++i) {
                    double Uadv = u(i, j, k) + u(i, j + 1, k) + u(i - 1, j, k) + u(i - 1, j + 1, k);
                    double Wadv = w(i, j, k) + w(i, j + 1, k) + w(i, j, k - 1) + w(i, j + 1, k - 1);
                    V2[ORC_IX(i, j, k)] = V2[ORC_IX(i, j, k)] -
                                          (Ax * (v(i + 1, j, k) - v(i - 1, j, k)) * Uadv +
                                           Ay * (v(i, j + 1, k) - v(i, j - 1, k)) * v(i, j, k) +
                                           Az * (v(i, j, k + 1) - v(i, j, k - 1)) * Wadv);
                }
    }
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) V2[ORC_IX(i, j, k)] = V2[ORC_IX(i, j, k)] * 0.5;
}

/* CDW: compiler/tests/fixtures/stencils/cdw.f90:12-125 (CDWdiv :44-71, CDWadv :73-98).
 * Note the operand order (V(i,j-1,k) + V(i,j-1,k+1)) in the y flux, kept as written.
 * Pinned by: bit equality with the reference cdw.f90 built into oracle/_ref and the 16^3 golden dump of
 * cdw_main.f90. */
void orc_cdw(double *W2, const double *U, const double *V, const double *W, double dxmin, double dymin,
             double dzmin, int nx, int ny, int nz, int e1, int e2, int e3)
{
    (void)e3;
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) W2[ORC_IX(i, j, k)] = 0.0;
    {
        const double Ax = 0.25 / dxmin, Ay = 0.25 / dymin, Az = 0.25 / dzmin;
#pragma omp parallel for
        for (int k = 2; k <= nz + 1; ++k)
            for (int j = 2; j <= ny + 1; ++j)
                for (int i = 2; i <= nx + 1; ++i) {
                    double z = Az * (w(i, j, k + 1) + w(i, j, k)) * (w(i, j, k + 1) + w(i, j, k)) -
                               Az * (w(i, j, k) + w(i, j, k - 1)) * (w(i, j, k) + w(i, j, k - 1));
                    double y = Ay * (w(i, j + 1, k) + w(i, j, k)) * (v(i, j, k + 1) + v(i, j, k)) -
                               Ay * (w(i, j, k) + w(i, j - 1, k)) * (v(i, j - 1, k) + v(i, j - 1, k + 1));
                    double x = Ax * (w(i + 1, j, k) + w(i, j, k)) * (u(i, j, k + 1) + u(i, j, k)) -
                               Ax * (w(i, j, k) + w(i - 1, j, k)) * (u(i - 1, j, k + 1) + u(i - 1, j, k));
                    W2[ORC_IX(i, j, k)] = -((z + y) + x);
                }
    }
    {
        const double Ax = 0.125 / dxmin, Ay = 0.125 / dymin, Az = 0.5 / dzmin;
#pragma omp parallel for
        for (int k = 2; k <= nz + 1; ++k)
            for (int j = 2; j <= ny + 1; ++j)
                for (int i = 2; i <= nx + 1; ++i) {
                    double Uadv = u(i, j, k) + u(i, j, k + 1) + u(i - 1, j, k) + u(i - 1, j, k + 1);
                    double Vadv = v(i, j, k) + v(i, j, k + 1) + v(i, j - 1, k) + v(i, j - 1, k + 1);
                    W2[ORC_IX(i, j, k)] = W2[ORC_IX(i, j, k)] -
                                          (Ax * (w(i + 1, j, k) - w(i - 1, j, k)) * Uadv +
                                           Ay * (w(i, j + 1, k) - w(i, j - 1, k)) * Vadv +
                                           Az * (w(i, j, k + 1) - w(i, j, k - 1)) * w(i, j, k));
                }
    }
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) W2[ORC_IX(i, j, k)] = W2[ORC_IX(i, j, k)] * 0.5;
}

/* Flux-form second-order centred scalar advection. The reference has no text for it (ELMM's
 * scalar_advection.f90 is not in the reference, benchmarks/elmm/elmm.toml:46); the nest is written in the
 * style of CDUdiv (cdu.f90:48-78): face velocity times the two-point mean of the scalar,
 *   s2 = -((Ax*U(i)*(s(i+1)+s(i)) - Ax*U(i-1)*(s(i)+s(i-1))) + (y) + (z)),  Ax = 0.5/dx.
 * Pinned by: exact zero for a uniform scalar in a discretely divergence-free field, discrete conservation
 * (sum of s2 is zero to rounding with periodic x,y and w = 0 on the walls), and a hand-computed value for
 * a linear scalar in a uniform wind (tests/test_oracle.py). */
void orc_scalar_adv(double *S2, const double *U, const double *V, const double *W, const double *S,
                    double dxmin, double dymin, double dzmin, int nx, int ny, int nz, int e1, int e2, int e3)
{
    (void)e3;
    const double Ax = 0.5 / dxmin, Ay = 0.5 / dymin, Az = 0.5 / dzmin;
#define s(a, b, c) S[ORC_IX(a, b, c)]
#pragma omp parallel for
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j)
            for (int i = 2; i <= nx + 1; ++i) {
                double x = Ax * u(i, j, k) * (s(i + 1, j, k) + s(i, j, k)) -
                           Ax * u(i - 1, j, k) * (s(i, j, k) + s(i - 1, j, k));
                double y = Ay * v(i, j, k) * (s(i, j + 1, k) + s(i, j, k)) -
                           Ay * v(i, j - 1, k) * (s(i, j, k) + s(i, j - 1, k));
                double z = Az * w(i, j, k) * (s(i, j, k + 1) + s(i, j, k)) -
                           Az * w(i, j, k - 1) * (s(i, j, k) + s(i, j, k - 1));
                S2[ORC_IX(i, j, k)] = -((x + y) + z);
            }
#undef s
#undef u
#undef v
#undef w
}
