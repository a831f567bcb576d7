/* ORACLE — test infrastructure only (see oracle.h).
 *
 * Ghost-layer fills.
 *
 * orc_box_periodic restates the box application's periodic halo (compiler/tests/apps/box/boundaries.f90:6-41):
 * two ghost layers, x faces first, then y faces over the whole x range, then z faces over the whole x-y
 * range, so edges and corners follow from the order. Pinned, with orc_box_diffuse_store, by bit equality
 * with the reference box application built natively into oracle/_ref/box: its field.bin after N steps, halos
 * included, equals N oracle steps applied to its field.bin after none
 * (tests/test_oracle.py::test_box_oracle_reproduces_the_reference_box_field_bit_for_bit).
 *
 * orc_halo_les fills the one-cell halo of the LES fields in the same x, y, z order: periodic in x and y;
 * at the z walls (ELMM's NOSLIP Btype, benchmarks/elmm/elmm.toml:194-200, whose source is not part of the
 * reference) the tangential velocities are mirrored with a sign change so that they vanish on the wall,
 * the wall-normal velocity is zero on and beyond the walls, and the scalar has zero gradient.
 * Pinned by brute force: every ghost equals its image (tests/test_oracle.py).
 */
#include "oracle.h"

enum { ORC_HALO_TANGENTIAL = 0, ORC_HALO_NORMAL = 1, ORC_HALO_SCALAR = 2 };

void orc_halo_les(double *F, int kind, int nx, int ny, int nz, int e1, int e2, int e3)
{
    (void)e3;
    for (int k = 2; k <= nz + 1; ++k)
        for (int j = 2; j <= ny + 1; ++j) {
            F[ORC_IX(1, j, k)] = F[ORC_IX(nx + 1, j, k)];
            F[ORC_IX(nx + 2, j, k)] = F[ORC_IX(2, j, k)];
        }
    for (int k = 2; k <= nz + 1; ++k)
        for (int i = 1; i <= nx + 2; ++i) {
            F[ORC_IX(i, 1, k)] = F[ORC_IX(i, ny + 1, k)];
            F[ORC_IX(i, ny + 2, k)] = F[ORC_IX(i, 2, k)];
        }
    for (int j = 1; j <= ny + 2; ++j)
        for (int i = 1; i <= nx + 2; ++i) {
            if (kind == ORC_HALO_TANGENTIAL) {
                F[ORC_IX(i, j, 1)] = -F[ORC_IX(i, j, 2)];
                F[ORC_IX(i, j, nz + 2)] = -F[ORC_IX(i, j, nz + 1)];
            } else if (kind == ORC_HALO_NORMAL) {
                F[ORC_IX(i, j, 1)] = 0.0;
                F[ORC_IX(i, j, nz + 1)] = 0.0;
                F[ORC_IX(i, j, nz + 2)] = 0.0;
            } else {
                F[ORC_IX(i, j, 1)] = F[ORC_IX(i, j, 2)];
                F[ORC_IX(i, j, nz + 2)] = F[ORC_IX(i, j, nz + 1)];
            }
        }
}

/* Box arrays are a(-1:nx+2, -1:ny+2, -1:nz+2). */
#define BX(i, j, k) ((size_t)((i) + 1) + (size_t)(nx + 4) * ((size_t)((j) + 1) + (size_t)(ny + 4) * (size_t)((k) + 1)))

void orc_box_periodic(double *a, int nx, int ny, int nz)
{
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= 2; ++i) {
                a[BX(nx + i, j, k)] = a[BX(i, j, k)];
                a[BX(i - 2, j, k)] = a[BX(nx + i - 2, j, k)];
            }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= 2; ++j)
            for (int i = -1; i <= nx + 2; ++i) {
                a[BX(i, ny + j, k)] = a[BX(i, j, k)];
                a[BX(i, j - 2, k)] = a[BX(i, ny + j - 2, k)];
            }
    for (int k = 1; k <= 2; ++k)
        for (int j = -1; j <= ny + 2; ++j)
            for (int i = -1; i <= nx + 2; ++i) {
                a[BX(i, j, nz + k)] = a[BX(i, j, k)];
                a[BX(i, j, k - 2)] = a[BX(i, j, nz + k - 2)];
            }
}

/* Box diffusion and store: compiler/tests/apps/box/stepper.f90:8-37.
 * b = a + dt*nu*laplacian, laplacian = (a(i-1)+a(i+1)+a(j-1)+a(j+1)+a(k-1)+a(k+1) - 6*a) / (dx*dx);
 * then the interior of a takes b. Pinned by the reference box application (oracle/_ref/box), as above. */
void orc_box_diffuse_store(double *a, double *b, int nx, int ny, int nz, double dx, double dt, double nu)
{
#pragma omp parallel for
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double lap = (a[BX(i - 1, j, k)] + a[BX(i + 1, j, k)] + a[BX(i, j - 1, k)] + a[BX(i, j + 1, k)] +
                              a[BX(i, j, k - 1)] + a[BX(i, j, k + 1)] - 6.0 * a[BX(i, j, k)]) /
                             (dx * dx);
                b[BX(i, j, k)] = a[BX(i, j, k)] + dt * nu * lap;
            }
#pragma omp parallel for
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) a[BX(i, j, k)] = b[BX(i, j, k)];
}
#undef BX
